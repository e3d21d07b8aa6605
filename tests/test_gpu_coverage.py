"""GPU parity of the generation corners: 2-D isotropic (both modes, FP64 and
FP32 per-realization directions), 1-D Shinozuka, the cosine correlation model
(cm_COS, spectra_RF.f90:210-220) and the BF16 shared-direction isotropic form
with the fused lognormal map, each against the FP64 oracle on the same points
and the same draws of the Fortran facade (random_number after random_seed).

Tolerances (relative RMS over the field vs FP64): FP64 1e-12, FP32 and MFMA
FP16 hi/lo 1e-5, MFMA BF16 hi/lo 4e-5 (stated looser bound).
"""
import math

import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = {R.FORM_FP64: 1e-12, R.FORM_FP32: 1e-5, R.FORM_MFMA_F16: 1e-5, R.FORM_MFMA_BF16: 4e-5}
NAMES = {R.FORM_FP64: "fp64", R.FORM_FP32: "fp32", R.FORM_MFMA_F16: "mfma-f16", R.FORM_MFMA_BF16: "mfma-bf16"}


def report(tag, a, b):
    e = float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))
    m = float(np.max(np.abs(a - b)))
    print(f"[parity] {tag}: rel_rms={e:.3e} max_abs={m:.3e}", flush=True)
    return e


@pytest.mark.parametrize("mode", [R.REF_COMPAT, R.FIXED])
@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32])
def test_isotropic_2d_forms(ctx, mode, form):
    """randomFieldND.f90:414-443: psi, theta, gamma per realization; REF_COMPAT
    theta = 2 pi psi and gamma = 2 pi U."""
    x = W.grid_points((40, 31), (0.1, 0.1), (2.0, -1.0))
    corrL = (0.5, 0.5)
    p = R.make_params(nDim=2, method=R.ISOTROPIC, mode=mode, form=form, Nmc=6, corrL=corrL,
                      xRange=(12.7, 12.7))
    _, U = R.reference_draws(p, R.reference_seed(7))
    po = O.make_plan(2, O.ISOTROPIC, mode, [25.4] * 2)
    assert U.shape == (6, 3 * po.kNTotal)
    k = O.kpoints(po)
    ref = O.isotropic(po, x / 0.5, k, O.skvec(po, k), U)
    got = R.generate(ctx, p, x, None, U)
    assert report(f"isotropic-2d mode={mode} {NAMES[form]}", got, ref) <= TOL[form]


@pytest.mark.parametrize("form", [R.FORM_FP32, R.FORM_MFMA_F16])
def test_isotropic_2d_shared_directions(ctx, form):
    x = W.grid_points((40, 31), (0.1, 0.1), (2.0, -1.0))
    p = R.make_params(nDim=2, method=R.ISOTROPIC, mode=R.FIXED, form=form, Nmc=6, corrL=(0.5, 0.5),
                      xRange=(12.7, 12.7), isoShared=1)
    _, U = R.reference_draws(p, R.reference_seed(7))
    kN = R.plan_spectrum(p).kNTotal
    Us = U.copy()
    Us[:, kN: 2 * kN] = U[0, kN: 2 * kN]          # 2-D: the theta row of realization 1
    po = O.make_plan(2, O.ISOTROPIC, O.FIXED, [25.4] * 2)
    k = O.kpoints(po)
    ref = O.isotropic(po, x / 0.5, k, O.skvec(po, k), Us)
    got = R.generate(ctx, p, x, None, U)
    assert report(f"isotropic-2d-shared {NAMES[form]}", got, ref) <= TOL[form]


@pytest.mark.parametrize("mode", [R.REF_COMPAT, R.FIXED])
@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16, R.FORM_MFMA_BF16])
def test_shinozuka_1d(ctx, mode, form):
    """1-D lattice: one sign block; REF_COMPAT leaves it unwritten (zero
    wavevectors, a constant field per realization, A1)."""
    x = np.sort(np.random.default_rng(3).random((5000, 1)) * 30.0, axis=0)
    p = R.make_params(nDim=1, method=R.SHINOZUKA, mode=mode, form=form, Nmc=5, corrL=(1.5,), xRange=(30.0,))
    _, phU = R.reference_draws(p, R.reference_seed(2))
    po = O.make_plan(1, O.SHINOZUKA, mode, [20.0])
    k = O.kpoints(po)
    ref = O.shinozuka(po, x / 1.5, k, O.skvec(po, k), phU)
    if mode == R.REF_COMPAT:
        assert np.allclose(ref, ref[:, :1], rtol=0, atol=1e-12)     # constant per realization
    got = R.generate(ctx, p, x, None, phU, dtype=np.float64 if form == R.FORM_FP64 else np.float32)
    assert report(f"shinozuka-1d mode={mode} {NAMES[form]}", got.astype(np.float64), ref) <= TOL[form]


@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16])
def test_cm_cos_single_wavevector(ctx, form):
    """cm_COS keeps S = 1 at the sixth wavevector only (spectra_RF.f90:210-220):
    the field is ampMult cos(k_6 . x + 2 pi U_6), a closed form, in FIXED mode."""
    x = W.cfg1_points(3000, 4)
    p = R.make_params(nDim=3, method=R.SHINOZUKA, corrMod=R.CM_COS, mode=R.FIXED, form=form, Nmc=3,
                      xRange=(10, 10, 10))
    _, phU = R.reference_draws(p, R.reference_seed(4))
    po = O.make_plan(3, O.SHINOZUKA, O.FIXED, [10, 10, 10], corrMod=2)
    k = O.kpoints(po)
    sk = O.skvec(po, k)
    assert np.count_nonzero(sk) == 1 and sk[5] == 1.0
    assert list(k[5]) == pytest.approx([5 * po.kDelta[0], 0.0, 0.0], abs=1e-15)
    ref = O.shinozuka(po, x, k, sk, phU)
    closed = po.ampMult * np.cos(x @ k[5] + 2 * O.PI_REF * phU[:, 5:6])
    assert np.allclose(ref, closed, rtol=0, atol=1e-12)
    got = R.generate(ctx, p, x, None, phU)
    assert report(f"cm_cos {NAMES[form]}", got, ref) <= TOL[form]


def test_cm_cos_ref_compat_constant(ctx):
    """REF_COMPAT: the sixth wavevector lies in the unwritten sign block 1
    (A1), so the field is the constant ampMult cos(2 pi U_6)."""
    x = W.cfg1_points(2000, 5)
    p = R.make_params(nDim=3, method=R.SHINOZUKA, corrMod=R.CM_COS, mode=R.REF_COMPAT, form=R.FORM_FP64,
                      Nmc=2, xRange=(10, 10, 10))
    _, phU = R.reference_draws(p, R.reference_seed(4))
    po = O.make_plan(3, O.SHINOZUKA, O.REF_COMPAT, [10, 10, 10], corrMod=2)
    k = O.kpoints(po)
    ref = O.shinozuka(po, x, k, O.skvec(po, k), phU)
    assert np.allclose(ref, po.ampMult * np.cos(2 * O.PI_REF * phU[:, 5:6]), rtol=0, atol=1e-13)
    got = R.generate(ctx, p, x, None, phU)
    assert report("cm_cos ref_compat fp64", got, ref) <= 1e-12


@pytest.mark.parametrize("form", [R.FORM_MFMA_BF16, R.FORM_MFMA_F16, R.FORM_FP32])
def test_cfg2_shape_lognormal_fused(ctx, form):
    """cfg 2's chain on a small grid: isotropic shared directions with the
    lognormal map fused into the epilogue (avg 1, var 0.25), FP32 output."""
    x = W.grid_points((16, 14, 12), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, form=form, Nmc=16, corrL=(0.5,) * 3,
                      xRange=(12.7,) * 3, isoShared=1, margiFirst=R.FOM_LOGNORMAL, fieldAvg=1.0, fieldVar=0.25)
    _, U = R.reference_draws(p, R.reference_seed(7))
    kN = R.plan_spectrum(p).kNTotal
    Us = U.copy()
    Us[:, : 2 * kN] = U[0, : 2 * kN]
    po = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    k = O.kpoints(po)
    ref = O.marginal(2, 1.0, 0.25, O.isotropic(po, x / 0.5, k, O.skvec(po, k), Us))
    got = R.generate(ctx, p, x, None, U, dtype=np.float32).astype(np.float64)
    assert report(f"cfg2-shape lognormal {NAMES[form]}", got, ref) <= TOL[form]


def test_grid_form_fp32_operand_fallback(ctx):
    """A first-contraction K beyond the FP16 operand range (2 kNStep_1 >
    16384: kNStep_1 = 8279 on a 6721-correlation-length axis) runs every stage
    on the FP32-operand tile; same result bar as the FP16 hi/lo tiles."""
    import torch
    nstep, step, corrL = (48, 3), (143.0, 1.0), (1.0, 1.0)
    ext = [(nstep[a] - 1) * step[a] for a in range(2)]
    p = R.make_params(nDim=2, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=3, corrL=corrL, xRange=ext)
    info = R.plan_spectrum(p)
    assert 2 * info.kNStep[0] > 16384
    _, phU = R.reference_draws(p, R.reference_seed(6))
    plan = R.FieldPlan(ctx, p, None, phU)
    out = torch.empty((3, int(np.prod(nstep))), dtype=torch.float32, device="cuda")
    plan.execute_grid((0.5, -1.0), step, nstep, out)
    ctx.sync()
    x = W.grid_points(nstep, step, (0.5, -1.0))
    po = O.make_plan(2, O.SHINOZUKA, O.FIXED, ext)
    k = O.kpoints(po)
    ref = O.shinozuka(po, x, k, O.skvec(po, k), phU)
    got = out.cpu().numpy().astype(np.float64)
    plan.close()
    assert report("grid form fp32-operand fallback", got, ref) <= 1e-5
