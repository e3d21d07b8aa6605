"""Edge cases of the point-wise generation path against the FP64 oracle:
empty and ragged point counts (not multiples of the 32/64/256-point blocks),
one realization and ragged realization counts, FP64 output of the MFMA form,
coordinates far from the origin (the FP64 reduction of k.x before the FP32
sines), and the C ABI's argument errors on device entry points."""
import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = {R.FORM_FP64: 1e-12, R.FORM_FP32: 1e-5, R.FORM_MFMA_F16: 1e-5}


def _case(Nmc, x, mode=R.FIXED):
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=mode, Nmc=Nmc, xRange=(6, 6, 6))
    _, phU = R.reference_draws(p, R.reference_seed(2))
    po = O.make_plan(3, O.SHINOZUKA, mode, [6, 6, 6])
    k = O.kpoints(po)
    ref = O.shinozuka(po, x, k, O.skvec(po, k), phU) if x.shape[0] else np.zeros((Nmc, 0))
    return p, phU, ref


def _run(ctx, p, phU, x, form, dt):
    import torch
    p.form = form
    plan = R.FieldPlan(ctx, p, None, phU)
    d_x = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    out = torch.full((p.Nmc, x.shape[0]), 7.0, dtype=dt, device="cuda")
    plan.execute(d_x, out)
    ctx.sync()
    plan.close()
    return out.cpu().numpy().astype(np.float64)


def _rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16])
@pytest.mark.parametrize("N", [1, 31, 33, 257, 1000])
def test_ragged_point_counts(ctx, form, N):
    import torch
    x = W.cfg1_points(N, 3) * 0.6
    p, phU, ref = _case(5, x)
    dt = torch.float64 if form == R.FORM_FP64 else torch.float32
    got = _run(ctx, p, phU, x, form, dt)
    print(f"[edge] N={N} form={form} rel={_rel(got, ref):.2e}", flush=True)
    assert got.shape == ref.shape and np.isfinite(got).all()
    assert _rel(got, ref) <= TOL[form]


@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16])
def test_empty_point_set(ctx, form):
    import torch
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=form, Nmc=3, xRange=(6, 6, 6), seed=1)
    plan = R.FieldPlan(ctx, p)
    out = torch.full((3, 4), 7.0, dtype=torch.float32, device="cuda")
    x = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    plan.execute(x, out[:, :0])
    ctx.sync()
    plan.close()
    assert (out.cpu().numpy() == 7.0).all()                # nothing written


@pytest.mark.parametrize("Nmc", [1, 33])
def test_realization_counts_mfma_fp64_out(ctx, Nmc):
    import torch
    x = W.cfg1_points(700, 4) * 0.6
    p, phU, ref = _case(Nmc, x)
    got = _run(ctx, p, phU, x, R.FORM_MFMA_F16, torch.float64)
    print(f"[edge] Nmc={Nmc} mfma fp64-out rel={_rel(got, ref):.2e}", flush=True)
    assert _rel(got, ref) <= 1e-5


def test_far_coordinates(ctx):
    """Points 1e5 correlation lengths from the origin: k.x ~ 1e5-1e6 radians,
    reduced in FP64 before the FP32 sines, so the FP32 and MFMA forms keep
    their tolerance (an FP32 argument would lose every significant bit)."""
    import torch
    x = W.cfg1_points(600, 5) * 0.6 + np.array([1.0e5, -2.5e5, 7.0e4])
    p, phU, ref = _case(4, x)
    for form in (R.FORM_FP32, R.FORM_MFMA_F16):
        got = _run(ctx, p, phU, x, form, torch.float32)
        print(f"[edge] far form={form} rel={_rel(got, ref):.2e}", flush=True)
        assert _rel(got, ref) <= 1e-5


def test_device_entry_argument_errors(ctx):
    import torch
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=2, xRange=(6, 6, 6), seed=1)
    plan = R.FieldPlan(ctx, p)
    import ctypes as C
    x = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    out = torch.zeros((2, 4), dtype=torch.float32, device="cuda")
    xp, op = C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr())
    L = R.lib()
    assert L.rf_plan_execute(plan._h, xp, 4, op, 7, None) == R.RF_E_ARG            # unknown output dtype
    assert L.rf_plan_execute(plan._h, None, 4, op, R.DTYPE_F32, None) == R.RF_E_ARG  # NULL points
    assert L.rf_plan_execute(plan._h, xp, -1, op, R.DTYPE_F32, None) == R.RF_E_ARG   # negative count
    assert L.rf_plan_execute(None, xp, 4, op, R.DTYPE_F32, None) == R.RF_E_ARG       # NULL plan
    assert L.rf_plan_execute_grid(plan._h, None, None, None, op, R.DTYPE_F32, None) == R.RF_E_ARG
    ctx.sync()
    plan.close()
