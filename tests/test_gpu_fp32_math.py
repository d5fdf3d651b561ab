"""The FP32-flux path's elementary functions on the device (``bt_fp32_math``)
against float64 NumPy: ``exp`` within 3 ulps on |x| <= 30 (the laws'
exponents: GHK |alpha| <= ~8, the pumps' Keq exponents ~20), ``expm1`` within
2 ulps on |x| <= 1/2 (down to the 1e-25 nonce scale), and the
reciprocal-and-Newton quotient within an ulp of the IEEE FP32 quotient.
These functions are this build's own (the reference evaluates its laws in
FP64), so the mathematics is the check: a wrong constant of log2(e), a
dropped correction term or a wrong Taylor coefficient fails these bounds.
"""

import ctypes
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run(fn, x, y=None):
    from betse_amd import _native
    lib = _native.load()
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(dev)
    yd = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float32)).to(dev) if y is not None else None
    out = torch.empty_like(xd)
    stream = torch.cuda.current_stream(dev)
    rc = lib.bt_fp32_math(ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(yd.data_ptr() if yd is not None else 0),
                          ctypes.c_void_p(out.data_ptr()), len(x), fn, ctypes.c_void_p(stream.cuda_stream))
    assert rc == _native.BT_OK
    torch.cuda.synchronize(dev)
    return out.cpu().numpy()


def ulps(got, ref):
    """|got - ref| in units of the FP32 spacing at ref."""
    ref32 = np.abs(ref).astype(np.float32)
    return np.abs(got.astype(np.float64) - ref) / np.spacing(ref32).astype(np.float64)


def test_fp32_exp():
    say("case: fp32 exp on [-30, 30]")
    rng = np.random.default_rng(11)
    x = np.concatenate([np.linspace(-30, 30, 200001), rng.uniform(-8, 8, 200000),
                        [0.0, 1e-30, -1e-30, 0.5, -0.5, 1.0, -1.0]]).astype(np.float32)
    got = run(0, x)
    err = ulps(got, np.exp(x.astype(np.float64)))
    say(f"exp32: max {err.max():.2f} ulp, mean {err.mean():.3f} ulp")
    assert err.max() <= 3.0


def test_fp32_expm1_small():
    say("case: fp32 expm1 on [-1/2, 1/2]")
    rng = np.random.default_rng(12)
    tiny = rng.uniform(-1, 1, 1000) * 10.0 ** rng.uniform(-25, -3, 1000)
    x = np.concatenate([np.linspace(-0.5, 0.5, 200001), tiny, [0.0, 1e-25, -1e-25]]).astype(np.float32)
    got = run(1, x)
    ref = np.expm1(x.astype(np.float64))
    nz = ref != 0
    err = ulps(got[nz], ref[nz])
    say(f"expm1_32: max {err.max():.2f} ulp, mean {err.mean():.3f} ulp")
    assert err.max() <= 2.0
    assert np.all(got[~nz] == 0)


def test_fp32_quotient():
    say("case: fp32 quotient over magnitudes 1e-18..1e18")
    rng = np.random.default_rng(13)
    n = 400000
    a = (rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-18, 18, n)).astype(np.float32)
    b = (rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-18, 18, n)).astype(np.float32)
    got = run(2, a, b)
    ieee = (a / b).astype(np.float32)                    # NumPy's FP32 division is IEEE
    err = np.abs(got.astype(np.float64) - ieee.astype(np.float64)) / np.spacing(np.abs(ieee)).astype(np.float64)
    say(f"dv32: max {err.max():.2f} ulp from IEEE, exact in {np.mean(got == ieee):.4f} of cases")
    assert err.max() <= 1.0
