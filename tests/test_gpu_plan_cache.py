"""rf_generate / rf_generate_grid keep their plans on the context: a call that
matches a kept plan (same parameter bytes, wavevector draws and mask) only
refreshes its draws.  Every call must equal, bit for bit, a plan created for
that call alone (the same kernels on the same inputs), through cache hits with
caller draws, with device-stream draws, after the staging buffer grows, after
eviction, and for keys that differ only in the wavevector draws or the mask."""
import numpy as np
import pytest
import torch

import randomfield_amd as R

pytestmark = pytest.mark.gpu


def _fresh(ctx, p, x, kU, phU, calc, dtype):
    plan = R.FieldPlan(ctx, p, kU, phU, calc)
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    out = torch.empty((p.Nmc, x.shape[0]), dtype=dtype, device="cuda")
    plan.execute(xd, out)
    torch.cuda.synchronize()
    plan.close()
    return out.cpu().numpy()


def _fresh_grid(ctx, p, step, gn, phU, calc):
    plan = R.FieldPlan(ctx, p, None, phU, calc)
    out = torch.empty((p.Nmc, int(np.prod(gn))), dtype=torch.float32, device="cuda")
    plan.execute_grid((0.0, 0.0, 0.0), step, gn, out)
    torch.cuda.synchronize()
    plan.close()
    return out.cpu().numpy()


def _draws(p, mode, seed, calc):
    """(kU, phU): the reference's draws in REF_COMPAT, plain uniforms in FIXED."""
    if mode == R.REF_COMPAT:
        return R.reference_draws(p, R.reference_seed(seed), calc)
    info = R.plan_spectrum(p)
    rng = np.random.default_rng(seed)
    kU = rng.random(info.nKUniform) if p.method == R.RANDOMIZATION else None
    return kU, rng.random((p.Nmc, info.nPhaseUniform))


CASES = [
    # nDim, method, mode, form, isoShared, Nmc, masked
    (3, R.SHINOZUKA, R.FIXED, R.FORM_MFMA_F16, 0, 4, False),
    (2, R.RANDOMIZATION, R.REF_COMPAT, R.FORM_FP64, 0, 3, True),
    (3, R.RANDOMIZATION, R.FIXED, R.FORM_FP32, 0, 2, False),
    (3, R.ISOTROPIC, R.FIXED, R.FORM_FP32, 1, 3, False),
    (2, R.ISOTROPIC, R.REF_COMPAT, R.FORM_FP64, 0, 2, False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d-m%d-mode%d-form%d-sh%d" % c[:5])
def test_cached_points_equal_fresh_plans(case):
    nDim, method, mode, form, shared, Nmc, masked = case
    print(f"[plan cache] points case={case}", flush=True)
    ctx = R.Context(0)
    xr = (3.0, 2.5, 2.0)[:nDim]
    p = R.make_params(nDim=nDim, method=method, mode=mode, form=form, Nmc=Nmc, xRange=xr, seed=77,
                      isoShared=shared)
    calc = np.array([1, 0, 1][:Nmc] + [1] * max(0, Nmc - 3), dtype=np.uint8) if masked else None
    rng = np.random.default_rng(3)
    small = rng.random((301, nDim)) * np.array(xr)
    large = rng.random((5003, nDim)) * np.array(xr)
    kU, ph1 = _draws(p, mode, 11, calc)
    _, ph2 = _draws(p, mode, 12, calc)
    dtype = torch.float64 if form == R.FORM_FP64 else torch.float32
    ndt = np.float64 if dtype == torch.float64 else np.float32
    # miss; hit with new caller draws; hit with the device stream; hit on a larger
    # staging buffer; back to the first draws on a smaller call
    for x, ph in ((small, ph1), (small, ph2), (small, None), (large, ph2), (small, ph1)):
        got = R.generate(ctx, p, x, kU, ph, calc, dtype=ndt)
        ref = _fresh(ctx, p, x, kU, ph, calc, dtype)
        assert np.array_equal(got, ref), (case, x.shape, ph is None)
    ctx.close()


def test_wavevector_draws_and_mask_are_part_of_the_key():
    print("[plan cache] key", flush=True)
    ctx = R.Context(0)
    p = R.make_params(nDim=2, method=R.RANDOMIZATION, mode=R.FIXED, form=R.FORM_FP64, Nmc=3,
                      xRange=(2.0, 2.0), seed=5)
    info = R.plan_spectrum(p)
    rng = np.random.default_rng(8)
    x = rng.random((400, 2)) * 2.0
    phU = rng.random((3, info.nPhaseUniform))
    kA, kB = rng.random(info.nKUniform), rng.random(info.nKUniform)
    mask = np.array([1, 0, 1], dtype=np.uint8)
    for kU, calc in ((kA, None), (kB, None), (kA, mask), (kA, None), (None, None), (kB, mask)):
        got = R.generate(ctx, p, x, kU, phU, calc)
        ref = _fresh(ctx, p, x, kU, phU, calc, torch.float64)
        assert np.array_equal(got, ref), (kU is kA, calc is not None)
    ctx.close()


def test_eviction_keeps_results():
    """Six parameter sets through a four-plan cache, then the first again."""
    print("[plan cache] eviction", flush=True)
    ctx = R.Context(0)
    rng = np.random.default_rng(9)
    x = rng.random((500, 3)) * 2.0
    plist = [R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_MFMA_F16, Nmc=2,
                           xRange=(2.0, 2.0, 2.0), seed=100 + i) for i in range(6)]
    for p in plist + plist[:2]:
        got = R.generate(ctx, p, x, dtype=np.float32)
        ref = _fresh(ctx, p, x, None, None, None, torch.float32)
        assert np.array_equal(got, ref), p.seed
    ctx.close()


@pytest.mark.parametrize("method,form", [(R.SHINOZUKA, R.FORM_MFMA_F16), (R.FFT, R.FORM_FP32)])
def test_cached_grid_equal_fresh_plans(method, form):
    print(f"[plan cache] grid method={method}", flush=True)
    ctx = R.Context(0)
    Nmc = 3
    if method == R.FFT:
        gn, step = (20, 18, 12), (0.1, 0.1, 0.1)
        xr = [(n - 1) * s for n, s in zip(gn, step)]
        p = R.make_params(nDim=3, method=method, mode=R.FIXED, form=form, Nmc=Nmc, xRange=xr, seed=21, gridN=gn)
    else:
        gn, step = (26, 19, 15), (0.2, 0.25, 0.3)
        xr = [(n - 1) * s for n, s in zip(gn, step)]
        p = R.make_params(nDim=3, method=method, mode=R.FIXED, form=form, Nmc=Nmc, xRange=xr, seed=21)
    info = R.plan_spectrum(p)
    rng = np.random.default_rng(4)
    ph = rng.random((Nmc, info.nPhaseUniform))
    for phU in (None, ph, None, ph):
        got = R.generate_grid(ctx, p, (0.0, 0.0, 0.0), step, gn, phU, dtype=np.float32)
        ref = _fresh_grid(ctx, p, step, gn, phU, None)
        assert np.array_equal(got, ref), (method, phU is None)
    ctx.close()
