"""rf_plan_redraw: a plan's fresh realizations (draw d) equal those of a plan
created with the same uniforms, realizations d*Nmc .. d*Nmc + Nmc - 1 of the
device phase stream, for every method, mode and form, on points and grids.
Equal inputs run the same kernels, so the outputs must agree bit for bit."""
import numpy as np
import pytest
import torch

import randomfield_amd as R

pytestmark = pytest.mark.gpu


def _stream_rows(ctx, seed, nPU, d):
    """Elements d*nPU .. (d+1)*nPU - 1 of stream (seed, tag 2), on the host."""
    buf = torch.empty((d + 1) * nPU, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, seed, 2, buf.numel(), buf)
    torch.cuda.synchronize()
    return buf[d * nPU:].cpu().numpy()


CASES = [
    # nDim, method, mode, form, isoShared, Nmc
    (3, R.SHINOZUKA, R.FIXED, R.FORM_FP64, 0, 3),
    (3, R.SHINOZUKA, R.FIXED, R.FORM_MFMA_F16, 0, 5),
    (2, R.SHINOZUKA, R.REF_COMPAT, R.FORM_FP64, 0, 2),
    (2, R.RANDOMIZATION, R.FIXED, R.FORM_FP32, 0, 4),
    (3, R.ISOTROPIC, R.FIXED, R.FORM_FP64, 0, 2),
    (3, R.ISOTROPIC, R.FIXED, R.FORM_MFMA_F16, 1, 3),
    (2, R.ISOTROPIC, R.REF_COMPAT, R.FORM_FP32, 0, 2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%d-m%d-mode%d-form%d-sh%d" % c[:5])
def test_redraw_points(case):
    nDim, method, mode, form, shared, Nmc = case
    seed = 1234 + method
    xr = (4.0, 3.0, 2.5)[:nDim]
    print(f"[redraw] points case={case}", flush=True)
    ctx = R.Context(0)
    p = R.make_params(nDim=nDim, method=method, mode=mode, form=form, Nmc=Nmc, xRange=xr, seed=seed,
                      isoShared=shared)
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.random((777, nDim)) * np.array(xr)).cuda()
    base = R.FieldPlan(ctx, p)
    nPU = base.info.nPhaseUniform * Nmc
    first = torch.empty((Nmc, x.shape[0]), dtype=torch.float64, device="cuda")
    base.execute(x, first)
    for d in (3, 0, 1):
        base.redraw(d)
        got = torch.empty_like(first)
        base.execute(x, got)
        ref_plan = R.FieldPlan(ctx, p, None, _stream_rows(ctx, seed, nPU, d).reshape(Nmc, -1))
        ref = torch.empty_like(first)
        ref_plan.execute(x, ref)
        ref_plan.close()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (case, d)
        if d == 0:
            assert torch.equal(got, first), case      # draw 0 is the creation draw
        else:
            assert not torch.equal(got, first), case
    base.close()
    ctx.close()


@pytest.mark.parametrize("method,form", [(R.SHINOZUKA, R.FORM_MFMA_F16), (R.FFT, R.FORM_FP32)])
def test_redraw_grid(method, form):
    """Grid executor: the folded coefficients are re-formed after a redraw."""
    Nmc, seed = 4, 99
    print(f"[redraw] grid method={method} form={form}", flush=True)
    ctx = R.Context(0)
    if method == R.FFT:
        gn, step = (24, 20, 16), (0.1, 0.1, 0.1)
        xr = [(n - 1) * s for n, s in zip(gn, step)]
        p = R.make_params(nDim=3, method=method, mode=R.FIXED, form=form, Nmc=Nmc, xRange=xr, seed=seed, gridN=gn)
    else:
        gn, step = (30, 22, 17), (0.2, 0.25, 0.3)
        xr = [(n - 1) * s for n, s in zip(gn, step)]
        p = R.make_params(nDim=3, method=method, mode=R.FIXED, form=form, Nmc=Nmc, xRange=xr, seed=seed)
    n = int(np.prod(gn))
    base = R.FieldPlan(ctx, p)
    nPU = base.info.nPhaseUniform * Nmc
    first = torch.empty((Nmc, n), dtype=torch.float32, device="cuda")
    base.execute_grid((0.0, 0.0, 0.0), step, gn, first)
    for d in (2, 5):
        base.redraw(d)
        got = torch.empty_like(first)
        base.execute_grid((0.0, 0.0, 0.0), step, gn, got)
        ref_plan = R.FieldPlan(ctx, p, None, _stream_rows(ctx, seed, nPU, d).reshape(Nmc, -1))
        ref = torch.empty_like(first)
        ref_plan.execute_grid((0.0, 0.0, 0.0), step, gn, ref)
        ref_plan.close()
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (method, d)
        assert not torch.equal(got, first)
    base.close()
    ctx.close()


def test_redraw_rejects_negative_draw():
    ctx = R.Context(0)
    plan = R.FieldPlan(ctx, R.make_params(nDim=2, Nmc=1, mode=R.FIXED, xRange=(2.0, 2.0)))
    with pytest.raises(R.RFError):
        plan.redraw(-1)
    plan.close()
    ctx.close()
