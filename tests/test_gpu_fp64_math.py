"""The FP64 path's own exponentials on the device (``bt_fp64_math``): the GHK
laws' ``exp`` (Cody-Waite reduction by ln 2, Taylor series to r^13, ldexp)
and ``expm1`` (Taylor series to x^17 on |x| <= 1/2), which replace the
library's functions in the membrane kernel (the reference calls NumPy's,
``sim_toolbox.py:18-69``). Checked against exact values (Python's decimal
arithmetic at 90 digits) on a sample, and against
NumPy (glibc, itself within an ulp) on a dense sweep: ``exp`` within 1.5 ulps
of the exact value on |x| <= 40 and on the normal range out to +-700,
``expm1`` within 1.5 ulps on |x| <= 1/2 down to the 1e-25 nonce scale; exp
overflows to +inf and NaN stays NaN. A wrong split of ln 2, a dropped Taylor
term or a coefficient off by an ulp fails these bounds. The quotient (two
Newton steps on the hardware reciprocal, one Markstein correction) against
IEEE division over magnitudes 1e-150..1e150 and the laws' own operand
ranges: at most an ulp apart, and equal in all but a few cases per million.
"""

import ctypes
import decimal
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
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)
    yd = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float64)).to(dev) if y is not None else None
    out = torch.empty_like(xd)
    stream = torch.cuda.current_stream(dev)
    rc = lib.bt_fp64_math(ctypes.c_void_p(xd.data_ptr()), ctypes.c_void_p(yd.data_ptr() if yd is not None else 0),
                          ctypes.c_void_p(out.data_ptr()), len(x), fn, ctypes.c_void_p(stream.cuda_stream))
    assert rc == _native.BT_OK
    torch.cuda.synchronize(dev)
    return out.cpu().numpy()


def exact(fn, x):
    """exp or expm1 of each FP64 x: 90 significant digits, so that exp(x) - 1
    keeps 40 of them down to |x| ~ 1e-30."""
    ctx = decimal.Context(prec=90)
    out = []
    for v in x:
        d = ctx.exp(decimal.Decimal(float(v)))
        out.append(ctx.subtract(d, 1) if fn == 1 else d)
    return out


def ulps_exact(got, ref_dec):
    """|got - ref| in units of the FP64 spacing at ref (ref at 40 digits)."""
    e = []
    for g, r in zip(got, ref_dec):
        sp = np.spacing(abs(float(r)))
        e.append(abs(float(decimal.Decimal(float(g)) - r)) / sp)
    return np.array(e)


def ulps(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_fp64_exp():
    rng = np.random.default_rng(5)
    dense = np.concatenate([np.linspace(-40, 40, 400001), rng.uniform(-700, 700, 100000),
                            [0.0, 1e-300, -1e-300, 1e-25, 0.5, -0.5, 1.0, -1.0, np.log(2.0), 700.0, -700.0]])
    got = run(0, dense)
    err = ulps(got, np.exp(dense))
    say(f"exp64 vs numpy: max {err.max():.2f} ulp, mean {err.mean():.3f} ulp")
    assert err.max() <= 2.5
    sample = np.concatenate([rng.uniform(-40, 40, 3000), rng.uniform(-700, 700, 1000),
                             rng.uniform(-1e-3, 1e-3, 500), [0.0, 1e-25, np.log(2.0) / 2, -np.log(2.0) / 2]])
    got_s = run(0, sample)
    e = ulps_exact(got_s, exact(0, sample))
    say(f"exp64 vs exact: max {e.max():.3f} ulp, mean {e.mean():.3f} ulp")
    assert e.max() <= 1.5
    special = run(0, np.array([800.0, np.nan, -800.0]))
    assert np.isposinf(special[0]) and np.isnan(special[1]) and special[2] == 0.0


def test_fp64_expm1_small():
    rng = np.random.default_rng(6)
    mags = 10.0 ** rng.uniform(-25, np.log10(0.5), 200000)
    dense = np.concatenate([np.linspace(-0.5, 0.5, 200001), mags, -mags, [0.0]])
    got = run(1, dense)
    nz = dense != 0
    err = ulps(got[nz], np.expm1(dense[nz]))
    say(f"expm1_64 vs numpy: max {err.max():.2f} ulp, mean {err.mean():.3f} ulp")
    assert err.max() <= 2.5 and got[~nz][0] == 0.0
    sample = np.concatenate([rng.uniform(-0.5, 0.5, 3000), 10.0 ** rng.uniform(-25, -1, 1000),
                             -(10.0 ** rng.uniform(-25, -1, 1000)), [0.5, -0.5, 1e-25]])
    got_s = run(1, sample)
    e = ulps_exact(got_s, exact(1, sample))
    say(f"expm1_64 vs exact: max {e.max():.3f} ulp, mean {e.mean():.3f} ulp")
    assert e.max() <= 1.5


def test_fp64_quotient():
    rng = np.random.default_rng(7)
    n = 1_000_000
    a = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-150, 150, n)
    b = rng.uniform(-1, 1, n) * 10.0 ** rng.uniform(-150, 150, n)
    # the GHK laws' quotients: concentrations over 1 - e^-alpha, |alpha| down to the nonce scale
    ca = 10.0 ** rng.uniform(-6, 3, n)
    al = np.where(rng.uniform(size=n) < 0.5, 1, -1) * 10.0 ** rng.uniform(-24, 1.2, n)
    x = np.concatenate([a, ca * (1 - np.exp(-al)) + ca, [1.0, 3.0, -7.0, 0.0]])
    y = np.concatenate([b, -np.expm1(-al), [3.0, 7.0, 3.0, 5.0]])
    got = run(2, x, y)
    ref = x / y
    nz = ref != 0
    err = ulps(got[nz], ref[nz])
    diff = int(np.count_nonzero(got != ref))
    say(f"quotient: max {err.max():.2f} ulp, {diff} of {len(x)} differ from IEEE")
    assert err.max() <= 1.0 and diff <= 5e-6 * len(x) and got[~nz][0] == 0.0

