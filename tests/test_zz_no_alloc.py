"""rf_normalize_device_noalloc (static device scratch, no allocation) against
rf_normalize_device on the same fields, through the C ABI: two calls in a row,
the second on a smaller field, FP64 and FP32; and the refusal of an Nmc above
the static scratch's limit.  Tolerance: relative 1e-13 of the field's scale
(the same kernels and sum partition; the coefficients are formed on the
device instead of the host)."""
import numpy as np
import pytest

import randomfield_amd as R

pytestmark = pytest.mark.gpu


def _pair(ctx, f, dtype_code):
    import torch
    a = torch.from_numpy(f.copy()).cuda()
    b = torch.from_numpy(f.copy()).cuda()
    Nmc, N = f.shape
    R.check(R.lib().rf_normalize_device(ctx.handle, R._ptr(a), N, Nmc, dtype_code, None), ctx.handle)
    R.check(R.lib().rf_normalize_device_noalloc(ctx.handle, R._ptr(b), N, Nmc, dtype_code, None), ctx.handle)
    ctx.sync()
    return a.cpu().numpy().astype(np.float64), b.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_matches_normalize(ctx, dtype):
    code = R.DTYPE_F64 if dtype == np.float64 else R.DTYPE_F32
    rng = np.random.default_rng(7)
    # first call: 300,007 points x 37 realizations; second: 1,001 x 5
    for N, Nmc in ((300007, 37), (1001, 5)):
        f = (rng.normal(3.0, 2.5, size=(Nmc, N))).astype(dtype)
        ref, got = _pair(ctx, f, code)
        scale = np.max(np.abs(ref))
        err = float(np.max(np.abs(got - ref)) / scale)
        same = bool(np.array_equal(got, ref))
        print(f"[noalloc] dtype={dtype.__name__} N={N} Nmc={Nmc} max_rel={err:.2e} bit_equal={same}", flush=True)
        tol = 1e-13 if dtype == np.float64 else 1e-6
        assert err <= tol
        m = got.mean(axis=1)
        s = np.sqrt((got ** 2).mean(axis=1))
        assert np.allclose(m, 0.0, atol=1e-6 if dtype == np.float32 else 1e-12)
        assert np.allclose(s, 1.0, rtol=1e-5 if dtype == np.float32 else 1e-12)


def test_noalloc_refuses_nmc_above_limit(ctx):
    import torch
    d = torch.zeros((2, 16), dtype=torch.float64, device="cuda")
    rc = R.lib().rf_normalize_device_noalloc(ctx.handle, R._ptr(d), 16, 1025, R.DTYPE_F64, None)
    assert rc == R.RF_E_ARG
    assert "1024" in R.lib().rf_last_error(ctx.handle).decode()
