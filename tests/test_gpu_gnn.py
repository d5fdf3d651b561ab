"""Warm-start network kernels on the GPU against the FP64 restatement.

The products run in BF16 with FP32 accumulation, so the bar is relative
(stated per test); the matrix product itself is checked against FP64 products
of the BF16-rounded operands, which it must match to FP32 accumulation error.
"""

import numpy as np
import pytest

from oracle import gnn_oracle as go

pytestmark = pytest.mark.gpu


def _bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from betse_amd.gnn import WarmStartGNN
    return torch, WarmStartGNN


@pytest.mark.parametrize("m,n,k,transA,transB,ksplit,beta", [
    (16, 16, 16, False, False, 1, 0),
    (37, 45, 70, False, False, 1, 0),
    (37, 45, 70, True, False, 1, 0),
    (37, 45, 70, False, True, 1, 1),
    (64, 64, 9000, True, False, 3, 0),
    (200, 8, 33, False, True, 1, 1),
])
def test_bf16_gemm(dev, m, n, k, transA, transB, ksplit, beta):
    torch, WarmStartGNN = dev
    print(f"gemm m={m} n={n} k={k} tA={transA} tB={transB} ksplit={ksplit} beta={beta}", flush=True)
    rng = np.random.RandomState(m * 7 + n + k)
    A = rng.randn(k, m) if transA else rng.randn(m, k)
    B = rng.randn(n, k) if transB else rng.randn(k, n)
    C0 = rng.randn(m, n)
    net = WarmStartGNN(4, 4, hidden=16, n_layers=1)
    At = torch.tensor(A.astype(np.float32), device=net.device)
    Bt = torch.tensor(B.astype(np.float32), device=net.device)
    Ct = torch.tensor(C0.astype(np.float32), device=net.device)
    sam, sak = (1, m) if transA else (k, 1)
    sbk, sbn = (1, k) if transB else (n, 1)
    ws = net._empty(max(ksplit, 1) * m * n + 16)
    rc = net.lib.bt_gnn_gemm(At.data_ptr(), sam, sak, Bt.data_ptr(), sbk, sbn, Ct.data_ptr(), n, m, n, k, beta,
                             ws.data_ptr(), ksplit, net.stream.cuda_stream)
    assert rc == 0
    net.sync()
    Aeff = _bf16(A.T if transA else A)
    Beff = _bf16(B.T if transB else B)
    want = Aeff @ Beff + (C0.astype(np.float32).astype(np.float64) if beta else 0.0)
    got = Ct.cpu().numpy().astype(np.float64)
    scale = np.abs(Aeff) @ np.abs(Beff) + 1.0
    err = np.max(np.abs(got - want) / scale)
    assert err < 1e-5, err


def _graph(n, rng, avg_deg=6):
    # ring-of-neighbours plus random chords, symmetric
    adj = [set() for _ in range(n)]
    for i in range(n):
        for d in (1, 2, 3):
            j = (i + d) % n
            adj[i].add(j)
            adj[j].add(i)
    for _ in range(n):
        i, j = rng.randint(0, n, 2)
        if i != j:
            adj[i].add(j)
            adj[j].add(i)
    ptr = np.concatenate([[0], np.cumsum([len(a) for a in adj])]).astype(np.int32)
    idx = np.concatenate([sorted(a) for a in adj]).astype(np.int32)
    return ptr, idx


def test_forward_backward_match_oracle(dev):
    torch, WarmStartGNN = dev
    rng = np.random.RandomState(11)
    n, fi, fo, H, L = 300, 10, 4, 64, 6
    ptr, idx = _graph(n, rng)
    X = rng.randn(n, fi)
    T = rng.randn(n, fo)
    mask = (rng.rand(n, fo) < 0.9).astype(np.float64)
    params = go.init_params(fi, fo, H, L, seed=2)
    for k in params:
        if k.startswith("b"):
            params[k] = rng.randn(*params[k].shape) * 0.05
    loss_o, g_o, y_o = go.loss_and_grads(params, X, T, mask, ptr.astype(np.int64), idx.astype(np.int64), L)
    net = WarmStartGNN(fi, fo, hidden=H, n_layers=L, params=params)
    d = net.device
    Xt = torch.tensor(X.astype(np.float32), device=d)
    Tt = torch.tensor(T.astype(np.float32), device=d)
    Mt = torch.tensor(mask.astype(np.float32), device=d)
    pt = torch.tensor(ptr, device=d)
    it = torch.tensor(idx, device=d)
    y = net.forward(Xt, pt, it)
    loss, dY = net.loss(y, Tt, Mt, float(mask.sum()))
    g = net.backward(dY, pt, it)
    net.sync()
    y_g = y.cpu().numpy().astype(np.float64)
    rel_y = np.linalg.norm(y_g - y_o) / np.linalg.norm(y_o)
    print("forward rel err", rel_y, "loss", loss, loss_o, flush=True)
    assert rel_y < 2e-2
    assert abs(loss - loss_o) / loss_o < 2e-2
    for name in go.param_names(L):
        a = g[name].cpu().numpy().astype(np.float64).reshape(g_o[name].shape)
        rel = np.linalg.norm(a - g_o[name]) / max(np.linalg.norm(g_o[name]), 1e-30)
        print(name, rel, flush=True)
        assert rel < 3e-2, (name, rel)


def test_adam_step_matches_numpy(dev):
    torch, WarmStartGNN = dev
    net = WarmStartGNN(3, 2, hidden=16, n_layers=1)
    rng = np.random.RandomState(4)
    p0 = {k: v.cpu().numpy().astype(np.float64) for k, v in net.params.items()}
    m = {k: np.zeros_like(v) for k, v in p0.items()}
    v2 = {k: np.zeros_like(v) for k, v in p0.items()}
    for t in (1, 2, 3):
        g = {k: rng.randn(*v.shape) for k, v in p0.items()}
        gt = {k: torch.tensor(v.astype(np.float32), device=net.device) for k, v in g.items()}
        net.adam(gt, lr=1e-2)
        for k in p0:
            m[k] = 0.9 * m[k] + 0.1 * g[k]
            v2[k] = 0.999 * v2[k] + 0.001 * g[k] ** 2
            p0[k] = p0[k] - 1e-2 * (m[k] / (1 - 0.9 ** t)) / (np.sqrt(v2[k] / (1 - 0.999 ** t)) + 1e-8)
    got = net.params_np()
    for k in p0:
        assert np.allclose(got[k], p0[k], rtol=1e-5, atol=1e-6), k


def test_training_reduces_loss(dev):
    torch, WarmStartGNN = dev
    rng = np.random.RandomState(12)
    n, fi, fo = 2000, 6, 3
    ptr, idx = _graph(n, rng)
    X = rng.randn(n, fi)
    # target: a smooth function of the features and of the neighbours' mean
    agg = go.mean_agg(X, ptr.astype(np.int64), idx.astype(np.int64))
    T = np.tanh(X[:, :fo] + 0.5 * agg[:, fo:2 * fo])
    net = WarmStartGNN(fi, fo, hidden=64, n_layers=3, seed=3)
    d = net.device
    Xt = torch.tensor(X.astype(np.float32), device=d)
    Tt = torch.tensor(T.astype(np.float32), device=d)
    Mt = torch.ones_like(Tt)
    pt, it = torch.tensor(ptr, device=d), torch.tensor(idx, device=d)
    losses = []
    for _ in range(150):
        y = net.forward(Xt, pt, it)
        loss, dY = net.loss(y, Tt, Mt, float(T.size))
        net.adam(net.backward(dY, pt, it), lr=3e-3)
        losses.append(loss)
    print("loss", losses[0], losses[-1], flush=True)
    assert losses[-1] < 0.1 * losses[0]
