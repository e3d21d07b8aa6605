"""The isotropic method on structured grids (rf_plan_execute_grid): each term
exp(i(k.x + phi)) factorises over the grid axes, so per realization the field
is the real part of one GEMM over the shells with per-axis phase tables as
operands (split-FP16 MFMA).  Checked against the FP64 oracle's point-by-point
isotropic sum (randomFieldND.f90:368-514 order) at the grid points, with the
reference's draws, in both modes, 2-D and 3-D, with shared directions, the
fused marginal maps and both output types."""
import numpy as np
import pytest
import torch

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _shared_rows(p, U):
    """The draws the shared-direction variant uses: realization 1's direction
    rows (phi, theta) for every realization."""
    kN = R.plan_spectrum(p).kNTotal
    Us = U.copy()
    Us[:, : 2 * kN] = U[0, : 2 * kN]
    return Us


CASES = [
    # nstep, step, origin, corrL, mode, isoShared, margi, out dtype
    ((20, 18, 16), (0.25, 0.25, 0.25), (1.5, -2.0, 0.25), (1.0, 1.2, 0.8), R.FIXED, 0, R.FOM_NONE, torch.float32),
    ((17, 9, 13), (0.3, 0.3, 0.3), (0.0, 0.0, 0.0), (0.7, 0.7, 0.7), R.REF_COMPAT, 0, R.FOM_NONE, torch.float32),
    ((130, 3, 5), (0.1, 0.5, 0.4), (0.0, 1.0, 2.0), (1.0, 1.0, 1.0), R.FIXED, 0, R.FOM_LOGNORMAL, torch.float64),
    ((12, 12, 12), (0.3, 0.3, 0.3), (0.0, 0.0, 0.0), (0.9, 0.9, 0.9), R.FIXED, 1, R.FOM_GAUSSIAN, torch.float32),
    ((33, 41), (0.2, 0.15), (0.5, -1.0), (1.0, 1.3), R.FIXED, 0, R.FOM_NONE, torch.float32),
    ((40, 29), (0.25, 0.25), (0.0, 0.0), (0.8, 0.8), R.REF_COMPAT, 0, R.FOM_NONE, torch.float64),
    # a last axis longer than 64 runs on 128-row tiles: one full, one partial
    ((9, 7, 150), (0.3, 0.4, 0.1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), R.FIXED, 0, R.FOM_NONE, torch.float32),
    ((50, 130), (0.2, 0.1), (1.0, 0.0), (1.0, 0.9), R.REF_COMPAT, 0, R.FOM_LOGNORMAL, torch.float32),
    # shell counts at the 16-shell chunk edges: 16 (one whole chunk, no
    # padding) and 6 (one chunk, 10 zero shells)
    ((11, 11, 11), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), R.FIXED, 0, R.FOM_NONE, torch.float32),
    ((6, 6), (0.1, 0.1), (0.0, 0.5), (1.0, 1.0), R.FIXED, 0, R.FOM_NONE, torch.float32),
]


@pytest.mark.parametrize("nstep,kn", [((11, 11, 11), 16), ((6, 6), 6)])
def test_iso_grid_chunk_edge_cases_have_those_shells(nstep, kn):
    """The two chunk-edge cases above keep the shell counts they are there for."""
    d = len(nstep)
    p = R.make_params(nDim=d, method=R.ISOTROPIC, mode=R.FIXED, form=R.FORM_FP32, Nmc=3, corrL=(1.0,) * d,
                      xRange=[(n - 1) * 0.1 for n in nstep])
    assert R.plan_spectrum(p).kNTotal == kn


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%s-mode%d-sh%d-m%d" % ("x".join(map(str, c[0])), c[4], c[5], c[6]))
def test_iso_grid_matches_oracle(ctx, case):
    nstep, step, origin, corrL, mode, shared, margi, dtype = case
    d = len(nstep)
    Nmc = 3
    ext = [(nstep[a] - 1) * step[a] for a in range(d)]
    print(f"[iso grid] case {case}", flush=True)
    p = R.make_params(nDim=d, method=R.ISOTROPIC, mode=mode, form=R.FORM_FP32, Nmc=Nmc, corrL=corrL, xRange=ext,
                      isoShared=shared, margiFirst=margi, fieldAvg=1.0, fieldVar=0.25)
    _, U = R.reference_draws(p, R.reference_seed(5))
    plan = R.FieldPlan(ctx, p, None, U)
    out = torch.empty((Nmc, int(np.prod(nstep))), dtype=dtype, device="cuda")
    plan.execute_grid(origin, step, nstep, out)
    ctx.sync()
    x = W.grid_points(nstep, step, origin)
    po = O.make_plan(d, O.ISOTROPIC, mode, [ext[a] / corrL[a] for a in range(d)])
    k = O.kpoints(po)
    ref = O.isotropic(po, x / np.array(corrL), k, O.skvec(po, k), _shared_rows(p, U) if shared else U)
    if margi != R.FOM_NONE:
        ref = O.marginal(margi, 1.0, 0.25, ref)
    got = out.cpu().numpy().astype(np.float64)
    e = rel_rms(got, ref)
    print(f"[iso grid] rel_rms={e:.3e} max_abs={np.max(np.abs(got - ref)):.3e}", flush=True)
    assert e <= 1e-5
    plan.close()


def test_iso_grid_redraw_and_masks(ctx):
    """After rf_plan_redraw the grid form uses the new directions, amplitudes
    and phases: equal, bit for bit, to a plan created with those draws; a
    masked realization gives the epilogue of zero, as the point forms do."""
    nstep, step = (24, 10, 9), (0.2, 0.3, 0.4)
    ext = [(nstep[a] - 1) * step[a] for a in range(3)]
    p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, form=R.FORM_FP32, Nmc=4, xRange=ext, seed=31)
    n = int(np.prod(nstep))
    base = R.FieldPlan(ctx, p)
    nPU = base.info.nPhaseUniform * 4
    buf = torch.empty(3 * nPU, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 31, 2, buf.numel(), buf)
    torch.cuda.synchronize()
    base.redraw(2)
    got = torch.empty((4, n), dtype=torch.float32, device="cuda")
    base.execute_grid((0.0, 0.0, 0.0), step, nstep, got)
    ref_plan = R.FieldPlan(ctx, p, None, buf[2 * nPU:].cpu().numpy().reshape(4, -1))
    ref = torch.empty_like(got)
    ref_plan.execute_grid((0.0, 0.0, 0.0), step, nstep, ref)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    calc = np.array([1, 0, 1, 1], dtype=np.uint8)
    masked = R.FieldPlan(ctx, p, None, None, calc)
    mo = torch.empty_like(got)
    masked.execute_grid((0.0, 0.0, 0.0), step, nstep, mo)
    pts = torch.from_numpy(W.grid_points(nstep, step, (0.0, 0.0, 0.0))).cuda()
    po = torch.empty_like(got)
    masked.execute(pts, po)
    torch.cuda.synchronize()
    assert torch.all(mo[1] == 0) and torch.all(po[1] == 0)
    assert rel_rms(mo.cpu().numpy().astype(np.float64), po.cpu().numpy().astype(np.float64)) <= 1e-5
    for pl in (base, ref_plan, masked):
        pl.close()
