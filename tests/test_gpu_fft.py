"""Method 4 (gen_Std_Gauss_FFT, randomFieldND.f90:670-735) on the device:
the grid form with DCT-III (REDFT01) tables against the FP64 oracle, at
rel RMS <= 1e-5 (FP32 tables, coefficients and MFMA accumulation)."""
import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _fft_case(ctx, gn, xs, corrL, Nmc, calc, seed):
    nDim = len(gn)
    step = [xs[a] * corrL[a] for a in range(nDim)]
    xr = [(gn[a] - 1) * step[a] for a in range(nDim)]
    p = R.make_params(nDim=nDim, method=R.FFT, mode=R.FIXED, form=R.FORM_FP32, Nmc=Nmc, corrL=corrL, xRange=xr,
                      gridN=gn)
    kN = int(np.prod(gn))
    u = np.random.default_rng(seed).random((Nmc, 3 * kN - 1))
    plan = R.FieldPlan(ctx, p, phU=u, calc=calc)
    po = O.fft_plan(nDim, [xr[a] / corrL[a] for a in range(nDim)], gn)
    ref = O.fft(po, O.skvec(po, O.kpoints(po)), u, xs, Nmc, calc)
    return plan, step, ref


def _rel_rms(got, ref):
    return float(np.sqrt(np.mean((got - ref) ** 2) / np.mean(ref ** 2)))


@pytest.mark.parametrize("gn,xs", [((24, 17), (0.2, 0.3)), ((32, 20, 12), (0.15, 0.2, 0.25)),
                                   ((64, 64, 64), (0.1, 0.1, 0.1))])
@pytest.mark.parametrize("odt", ["f32", "f64"])
def test_fft_grid_matches_oracle(ctx, gn, xs, odt):
    import torch
    corrL = (0.5, 1.0, 2.0)[:len(gn)]
    Nmc, calc = 3, [1, 0, 1]
    print(f"[fft] grid {gn} out {odt}", flush=True)
    plan, step, ref = _fft_case(ctx, gn, xs, corrL, Nmc, calc, sum(gn))
    out = torch.full((Nmc, int(np.prod(gn))), 7.0, dtype=torch.float32 if odt == "f32" else torch.float64,
                     device="cuda")
    plan.execute_grid([0.0] * len(gn), step, gn, out)
    ctx.sync()
    got = out.cpu().numpy().astype(np.float64)
    plan.close()
    err = _rel_rms(got[[0, 2]], ref[[0, 2]])
    print(f"[fft] rel RMS {err:.2e}", flush=True)
    assert err <= 1e-5
    assert not got[1].any()


def test_fft_errors_and_batch(ctx):
    """Point sets are refused (the reference transforms whole grids), grid
    counts must be the DCT sizes, and a grouped batch holding an FFT field
    next to a Shinozuka field gives the bits of the single launches."""
    import torch
    gn, xs, corrL = (20, 14, 10), (0.2, 0.25, 0.3), (1.0, 1.0, 1.0)
    plan, step, ref = _fft_case(ctx, gn, xs, corrL, 2, None, 5)
    kN = int(np.prod(gn))
    x = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    o = torch.empty((2, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(R.RFError) as e:
        plan.execute(x, o)
    assert e.value.code == R.RF_E_UNSUPPORTED
    bad = torch.empty((2, 21 * 14 * 10), dtype=torch.float32, device="cuda")
    with pytest.raises(R.RFError) as e:
        plan.execute_grid([0.0] * 3, step, (21, 14, 10), bad)
    assert e.value.code == R.RF_E_ARG
    sp = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=2, xRange=(4, 4, 4), seed=3)
    shin = R.FieldPlan(ctx, sp)
    outs = [torch.empty((2, kN), dtype=torch.float32, device="cuda") for _ in range(2)]
    geo = [((0.0, 0.0, 0.0), step, gn), ((0.1, 0.2, 0.3), (0.2, 0.2, 0.2), gn)]
    singles = []
    for pl, ob, (org, st, ns) in zip((plan, shin), outs, geo):
        pl.execute_grid(org, st, ns, ob)
        ctx.sync()
        singles.append(ob.cpu().numpy().copy())
        ob.zero_()
    assert _rel_rms(singles[0].astype(np.float64), ref) <= 1e-5
    b = R.GridBatch(ctx, [plan, shin], [g[0] for g in geo], [g[1] for g in geo], [g[2] for g in geo], outs,
                    R.DTYPE_F32)
    b.execute()
    ctx.sync()
    for i, ob in enumerate(outs):
        assert np.array_equal(ob.cpu().numpy(), singles[i]), i
    b.close()
    plan.close()
    shin.close()
