"""bt_gnn_colsum_noalloc, the entry point that allocates nothing, against
bt_gnn_colsum with a caller workspace of the same size: bit-identical column
sums on two calls in a row (the second smaller), on a call whose partials
need more rows per block, and the size limit."""

import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_colsum_noalloc_matches_the_workspace_entry_point():
    import torch
    from betse_amd import _native
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _native.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.Stream(device=dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    ws = torch.empty(_native.BT_GNN_SCRATCH_FLOATS, dtype=torch.float32, device=dev)
    rng = np.random.RandomState(5)
    for n, f, ldg in ((10_000, 64, 80), (3_000, 48, 48), (300_000, 256, 256), (1, 1, 1)):
        say(f"case: colsum_noalloc n={n} f={f} ldg={ldg}")
        g = rng.randn(n, ldg).astype(np.float32)
        with torch.cuda.stream(stream):
            G = torch.from_numpy(g).to(dev)
            a = torch.full((f,), float("nan"), dtype=torch.float32, device=dev)
            b = torch.full((f,), float("nan"), dtype=torch.float32, device=dev)
        stream.synchronize()
        rc = lib.bt_gnn_colsum(ctypes.c_void_p(G.data_ptr()), ldg, n, f, ctypes.c_void_p(a.data_ptr()),
                               ctypes.c_void_p(ws.data_ptr()), _native.BT_GNN_SCRATCH_FLOATS, sp)
        assert rc == _native.BT_OK
        rc = lib.bt_gnn_colsum_noalloc(ctypes.c_void_p(G.data_ptr()), ldg, n, f, ctypes.c_void_p(b.data_ptr()), sp)
        assert rc == _native.BT_OK
        stream.synchronize()
        ra, rb = a.cpu().numpy(), b.cpu().numpy()
        assert np.array_equal(ra, rb), float(np.abs(ra - rb).max())
        ref = g[:, :f].astype(np.float64).sum(axis=0)
        assert np.allclose(rb, ref, rtol=1e-4, atol=1e-3 * np.sqrt(n))
    # beyond the maximum width: rejected before any launch
    assert lib.bt_gnn_colsum_noalloc(ctypes.c_void_p(G.data_ptr()), 512, 4, 300,
                                     ctypes.c_void_p(b.data_ptr()), sp) == _native.BT_EINVAL
    assert lib.bt_gnn_colsum_noalloc(None, 1, 1, 1, ctypes.c_void_p(b.data_ptr()), sp) == _native.BT_EINVAL
