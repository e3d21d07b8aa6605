"""GPU parity: every precision form through the C ABI against the FP64 oracle,
on the same points and the same uniforms (drawn by the Fortran facade with the
compiler's random_number, as the reference draws them).

Tolerances (relative RMS error over the whole field, vs FP64):
  FP64 form          1e-12   (same operation order; device vs libm cos)
  FP32 form          1e-5
  MFMA FP16 hi/lo    1e-5
  MFMA BF16 hi/lo    4e-5    (stated looser bound: 8-bit mantissa halves)
"""
import math
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd import workloads as W

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent
TOL = {R.FORM_FP64: 1e-12, R.FORM_FP32: 1e-5, R.FORM_MFMA_F16: 1e-5, R.FORM_MFMA_BF16: 4e-5}
NAMES = {R.FORM_FP64: "fp64", R.FORM_FP32: "fp32", R.FORM_MFMA_F16: "mfma-f16", R.FORM_MFMA_BF16: "mfma-bf16"}


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def report(tag, a, b):
    e = rel_rms(a, b)
    m = float(np.max(np.abs(a - b)))
    print(f"[parity] {tag}: rel_rms={e:.3e} max_abs={m:.3e} rms_ref={np.sqrt(np.mean(b**2)):.3e}", flush=True)
    return e


# ------------------------------------------------------------- cfg1 shape
@pytest.fixture(scope="module")
def cfg1():
    x = W.cfg1_points(20000, 1)
    out = {}
    for mode in (R.REF_COMPAT, R.FIXED):
        p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=mode, Nmc=4, xRange=(10, 10, 10))
        _, phU = R.reference_draws(p, R.reference_seed(1))
        po = O.make_plan(3, O.SHINOZUKA, mode, [10, 10, 10])
        k = O.kpoints(po)
        ref = O.shinozuka(po, x, k, O.skvec(po, k), phU)
        out[mode] = (p, phU, ref)
    return x, out


@pytest.mark.parametrize("mode", [R.REF_COMPAT, R.FIXED])
@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16, R.FORM_MFMA_BF16])
def test_cfg1_forms(ctx, cfg1, mode, form):
    x, d = cfg1
    p, phU, ref = d[mode]
    p.form = form
    dt = np.float64 if form == R.FORM_FP64 else np.float32
    got = R.generate(ctx, p, x, None, phU, None, dtype=dt).astype(np.float64)
    e = report(f"cfg1 mode={mode} {NAMES[form]}", got, ref)
    assert e <= TOL[form]
    if form == R.FORM_FP64:
        assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_cfg1_ref_compat_structure(cfg1):
    # the oracle's REF_COMPAT realization n carries the constant block-1 offset
    # ampMult * sum cos(phi) and S^(2^-n): the spatial variance grows with n
    x, d = cfg1
    ref = d[R.REF_COMPAT][2]
    v = ref.var(axis=1)
    assert np.all(np.diff(v) > 0)


def test_calculate_mask_and_slabs(ctx):
    x = W.cfg1_points(3000, 5)
    calc = np.array([1, 0, 1, 1], np.uint8)
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.REF_COMPAT, Nmc=4, xRange=(10, 10, 10), xSlab=700)
    _, phU = R.reference_draws(p, [123], calc)
    po = O.make_plan(3, O.SHINOZUKA, O.REF_COMPAT, [10, 10, 10])
    k = O.kpoints(po)
    ref = O.shinozuka(po, x, k, O.skvec(po, k), phU, calc=calc, xSlab=700)
    for form in (R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16):
        p.form = form
        got = R.generate(ctx, p, x, None, phU, calc)
        assert not got[1].any()
        assert report(f"mask+slabs {NAMES[form]}", got, ref) <= TOL[form]


# ------------------------------------------------------ randomization (cfg3 shape)
@pytest.fixture(scope="module")
def cfg3_small():
    x = W.gaussian_mixture_points(4000, 11)
    corrL = (2.0, 2.0, 0.5)
    p = R.make_params(nDim=3, method=R.RANDOMIZATION, mode=R.FIXED, kSampling=R.KSAMP_DENSITY,
                      nKDensity=4096, Nmc=64, corrL=corrL, xRange=(1, 1, 1))
    kU, phU = R.reference_draws(p, R.reference_seed(11))
    po = O.make_plan(3, O.RANDOMIZATION, O.FIXED, [1, 1, 1], ksamp=O.KDENSITY, nKDensity=4096)
    k = O.kpoints(po, kU)
    ref = O.shinozuka(po, x / np.array(corrL), k, O.skvec(po, k), phU)
    return x, p, kU, phU, ref


@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16, R.FORM_MFMA_BF16])
def test_cfg3_small_forms(ctx, cfg3_small, form):
    x, p, kU, phU, ref = cfg3_small
    p.form = form
    got = R.generate(ctx, p, x, kU, phU, dtype=np.float64 if form == R.FORM_FP64 else np.float32)
    assert report(f"cfg3-small {NAMES[form]}", got.astype(np.float64), ref) <= TOL[form]


@pytest.mark.parametrize("Nmc", [64, 3, 300])
def test_mfma_variants_identical(ctx, cfg3_small, Nmc):
    """Every MFMA launch shape gives bit-identical fields (same per-point
    operation order); the autotuner keeps one of them and leaves its result."""
    import torch
    x, p0, kU, phU, ref = cfg3_small
    p = R.make_params(nDim=3, method=R.RANDOMIZATION, mode=R.FIXED, kSampling=R.KSAMP_DENSITY,
                      nKDensity=4096, Nmc=Nmc, corrL=(2.0, 2.0, 0.5), xRange=(1, 1, 1),
                      form=R.FORM_MFMA_F16)
    if Nmc != 64:
        kU, phU = R.reference_draws(p, R.reference_seed(11))
    plan = R.FieldPlan(ctx, p, kU, phU)
    dx = torch.from_numpy(x).cuda()
    outs = []
    for v in range(R.MFMA_VARIANTS):
        plan.set_variant(v)
        o = torch.empty((Nmc, x.shape[0]), dtype=torch.float32, device="cuda")
        plan.execute(dx, o)
        ctx.sync()
        outs.append(o.cpu().numpy())
    for v in range(1, R.MFMA_VARIANTS):
        assert np.array_equal(outs[v], outs[0]), v
    o = torch.empty((Nmc, x.shape[0]), dtype=torch.float32, device="cuda")
    chosen = plan.autotune(dx, o)
    ctx.sync()
    assert 0 <= chosen < R.MFMA_VARIANTS
    assert np.array_equal(o.cpu().numpy(), outs[0])
    with pytest.raises(R.RFError):
        plan.set_variant(R.MFMA_VARIANTS)
    plan.close()
    if Nmc == 64:
        assert report("mfma variants", outs[0].astype(np.float64), ref) <= TOL[R.FORM_MFMA_F16]


def test_randomization_box_ref_compat(ctx):
    x = W.cfg1_points(1500, 9)[:, :2] * 0.8
    p = R.make_params(nDim=2, method=R.RANDOMIZATION, mode=R.REF_COMPAT, Nmc=3, corrL=(1.0, 2.0),
                      xRange=(8, 8))
    kU, phU = R.reference_draws(p, [405])
    po = O.make_plan(2, O.RANDOMIZATION, O.REF_COMPAT, [8, 4])
    k = O.kpoints(po, kU)
    ref = O.shinozuka(po, x / np.array([1.0, 2.0]), k, O.skvec(po, k), phU)
    for form in (R.FORM_FP64, R.FORM_FP32, R.FORM_MFMA_F16):
        p.form = form
        got = R.generate(ctx, p, x, kU, phU)
        assert report(f"randomization-box {NAMES[form]}", got, ref) <= TOL[form]


# ------------------------------------------------------ isotropic (cfg2 lattice)
@pytest.fixture(scope="module")
def iso_case():
    x = W.grid_points((12, 10, 9), (0.1, 0.1, 0.1), (2.0, 3.0, 1.0))
    corrL = (0.5, 0.5, 0.5)
    return x, corrL


@pytest.mark.parametrize("mode", [R.REF_COMPAT, R.FIXED])
@pytest.mark.parametrize("form", [R.FORM_FP64, R.FORM_FP32])
def test_isotropic_forms(ctx, iso_case, mode, form):
    x, corrL = iso_case
    p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=mode, form=form, Nmc=8, corrL=corrL,
                      xRange=(12.7, 12.7, 12.7))
    _, U = R.reference_draws(p, R.reference_seed(7))
    po = O.make_plan(3, O.ISOTROPIC, mode, [25.4] * 3)
    k = O.kpoints(po)
    ref = O.isotropic(po, x / 0.5, k, O.skvec(po, k), U)
    got = R.generate(ctx, p, x, None, U)
    assert report(f"isotropic mode={mode} {NAMES[form]}", got, ref) <= TOL[form]


@pytest.mark.parametrize("form", [R.FORM_FP32, R.FORM_MFMA_F16, R.FORM_MFMA_BF16])
def test_isotropic_shared_directions(ctx, iso_case, form):
    x, corrL = iso_case
    p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, form=form, Nmc=8, corrL=corrL,
                      xRange=(12.7, 12.7, 12.7), isoShared=1)
    _, U = R.reference_draws(p, R.reference_seed(7))
    kN = R.plan_spectrum(p).kNTotal
    Us = U.copy()
    Us[:, : 2 * kN] = U[0, : 2 * kN]            # phi and theta rows of realization 1
    po = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    k = O.kpoints(po)
    ref = O.isotropic(po, x / 0.5, k, O.skvec(po, k), Us)
    got = R.generate(ctx, p, x, None, U)
    assert report(f"isotropic-shared {NAMES[form]}", got, ref) <= TOL[form]


def test_lognormal_fused(ctx, iso_case):
    x, corrL = iso_case
    p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, form=R.FORM_FP32, Nmc=4, corrL=corrL,
                      xRange=(12.7, 12.7, 12.7), margiFirst=R.FOM_LOGNORMAL, fieldAvg=1.0, fieldVar=0.25)
    _, U = R.reference_draws(p, R.reference_seed(7))
    po = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    k = O.kpoints(po)
    ref = O.marginal(2, 1.0, 0.25, O.isotropic(po, x / 0.5, k, O.skvec(po, k), U))
    got = R.generate(ctx, p, x, None, U, dtype=np.float32).astype(np.float64)
    assert report("lognormal fused fp32", got, ref) <= 1e-5
    # standalone transform on a host buffer
    g = np.random.default_rng(1).standard_normal((3, 1000))
    h = g.copy()
    R.multivariate_transformation(ctx, R.FOM_LOGNORMAL, 1.0, 0.25, h)
    assert np.allclose(h, O.marginal(2, 1.0, 0.25, g), rtol=1e-14)
    with pytest.raises(R.RFError) as e:
        R.multivariate_transformation(ctx, R.FOM_LOGNORMAL, -1.0, 0.25, h)
    assert e.value.code == R.RF_E_LOGNORMAL


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("offset,count", [(0, 100003), (1, 100003), (3, 7), (0, 4)])
def test_marginal_device_vector_and_tail(ctx, dtype, offset, count):
    """Aligned vector body + scalar tail, and a misaligned start (offset in
    elements), against the oracle's map of the same values."""
    import torch
    tt = torch.float32 if dtype == np.float32 else torch.float64
    g = np.random.default_rng(count + offset).standard_normal(offset + count).astype(dtype)
    d = torch.from_numpy(g.copy()).cuda()
    view = d[offset:]
    for margi, avg, var in ((R.FOM_LOGNORMAL, 1.0, 0.25), (R.FOM_GAUSSIAN, 2.0, 9.0)):
        view.copy_(torch.from_numpy(g[offset:]).to(tt))
        R.check(R.lib().rf_marginal_device(ctx.handle, margi, avg, var, R._ptr(view), count,
                                           R.DTYPE_F32 if dtype == np.float32 else R.DTYPE_F64, None))
        ctx.sync()
        ref = O.marginal(margi, avg, var, g[offset:].astype(np.float64)[None, :])[0].astype(dtype)
        got = d.cpu().numpy()
        assert np.array_equal(got[:offset], g[:offset])          # untouched before the view
        rt = 1e-6 if dtype == np.float32 else 1e-14
        assert np.allclose(got[offset:], ref, rtol=rt, atol=0), (margi, offset, count)


# ------------------------------------------------------------ device helpers
def test_device_uniforms_match_host(ctx):
    import torch
    d = torch.empty(5000, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 42, 2, 5000, d)
    ctx.sync()
    host = np.array([R.lib().rf_uniform_host(42, 2, e) for e in range(5000)])
    assert np.array_equal(d.cpu().numpy(), host)


@pytest.mark.parametrize("separable", [True, False])
def test_blend_matches_oracle(ctx, separable):
    """Both blend kernels: the separable one (masks a function of the
    intervals, as every tensor layout has) and the general cell-list one
    (one subdomain's mask changed so the per-axis form does not apply)."""
    import torch
    groupN = (40, 33, 21)
    origins, sizes, masks = W.cfg4_layout(groupN, (3, 2, 2), 5)
    rng = np.random.default_rng(3)
    Nmc = 3
    fields = [rng.standard_normal((Nmc, int(np.prod(s)))) for s in sizes]
    for mode in (R.REF_COMPAT, R.FIXED):
        fm = np.full(len(sizes), 63, np.int32) if mode == R.REF_COMPAT else np.array(masks, np.int32)
        if not separable:
            fm = fm.copy()
            fm[0] ^= 0b10          # upper x face of subdomain 0 only
        ref = O.blend(groupN, Nmc, origins, sizes, [5, 5, 5], fm, mode, fields)
        dfs = [torch.from_numpy(f).cuda() for f in fields]
        g = torch.empty((Nmc, int(np.prod(groupN))), dtype=torch.float64, device="cuda")
        R.blend(ctx, groupN, Nmc, origins, sizes, [5, 5, 5], fm, mode, dfs, g, R.DTYPE_F64)
        got = g.cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-15 * np.max(np.abs(ref)) * 10, mode


def test_stats_and_normalize(ctx):
    import torch
    f = np.random.default_rng(4).normal(2.0, 3.0, size=(5, 100003))
    d = torch.from_numpy(f).cuda()
    avg, sd, g = R.stats(ctx, d, f.shape[1], f.shape[0], R.DTYPE_F64)
    a, s, ga, gs = O.stats(f)
    assert np.allclose(avg, a, rtol=1e-12) and np.allclose(sd, s, rtol=1e-12)
    assert g[0] == pytest.approx(ga, rel=1e-12) and g[1] == pytest.approx(gs, rel=1e-12)
    R.normalize(ctx, d, f.shape[1], f.shape[0], R.DTYPE_F64)
    assert np.allclose(d.cpu().numpy(), O.normalize(f), rtol=1e-11, atol=1e-12)


# ------------------------------------------------------------ Fortran host
@pytest.mark.parametrize("case", [1, 2, 3, 4, 5, 6])
def test_fortran_driver(tmp_path, case):
    exe = ROOT / "tests" / "fortran" / "bin" / "rf_driver"
    out = tmp_path / f"case{case}.bin"
    r = subprocess.run([str(exe), str(case), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    N, Nmc, nDim = np.frombuffer(raw[:12], np.int32)
    kN = int(np.frombuffer(raw[12:20], np.int64)[0])
    off = 20
    x = np.frombuffer(raw[off: off + 8 * N * nDim], np.float64).reshape(N, nDim)
    off += 8 * N * nDim
    field = np.frombuffer(raw[off: off + 8 * N * Nmc], np.float64).reshape(Nmc, N)
    off += 8 * N * Nmc
    if case == 5:
        corrL = np.array([1.0, 1.5, 0.8])
        p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, Nmc=Nmc, corrL=corrL,
                          xRange=tuple(x.max(0) - x.min(0)))
        _, phU = R.reference_draws(p, R.reference_seed(3))
        po = O.make_plan(3, O.SHINOZUKA, O.FIXED, list((x.max(0) - x.min(0)) / corrL))
        k = O.kpoints(po)
        ref = O.marginal(2, 1.0, 0.25, O.shinozuka(po, x / corrL, k, O.skvec(po, k), phU))
        assert report("fortran createStandardGaussianFieldUnstruct", field, ref) <= 1e-12
        return
    if case == 6:
        # method 4 through createRandomField on the structured mesh: draws of
        # the facade (random_number after the seed of seedStart 5), the oracle's
        # DCT on the normalised steps
        gn, step, corrL = (12, 10, 8), np.array([0.25, 0.3, 0.4]), np.array([1.0, 1.5, 0.8])
        xr = (np.array(gn) - 1) * step
        p = R.make_params(nDim=3, method=R.FFT, mode=R.FIXED, Nmc=Nmc, corrL=corrL, xRange=tuple(xr), gridN=gn)
        _, U = R.reference_draws(p, R.reference_seed(5))
        po = O.fft_plan(3, list(xr / corrL), gn)
        kt = O.kpoints(po)
        kP = np.frombuffer(raw[off: off + 8 * 3 * kN], np.float64).reshape(kN, 3)
        assert kN == int(np.prod(gn)) and np.array_equal(kP, kt)
        ref = O.fft(po, O.skvec(po, kt), U, list(step / corrL), Nmc)
        assert report("fortran driver case 6 (method 4)", field, ref) <= 1e-5
        return
    rows = 1 if case == 4 else nDim
    kP = np.frombuffer(raw[off: off + 8 * rows * kN], np.float64).reshape(kN, rows)
    off += 8 * rows * kN
    sk = np.frombuffer(raw[off: off + 8 * kN], np.float64)
    if case in (1, 2):
        mode = O.REF_COMPAT if case == 1 else O.FIXED
        po = O.make_plan(3, O.SHINOZUKA, mode, [10, 10, 10])
        p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=mode, Nmc=Nmc, xRange=(10, 10, 10))
        _, phU = R.reference_draws(p, R.reference_seed(1))
        k = O.kpoints(po)
        assert np.array_equal(kP, k) and np.array_equal(sk, O.skvec(po, k))
        ref = O.shinozuka(po, x, k, sk, phU)
        tol = 1e-12 if case == 1 else 1e-5
    elif case == 3:
        corrL = np.array([1.0, 2.0])
        calc = np.array([1, 0], np.uint8)
        p = R.make_params(nDim=2, method=R.RANDOMIZATION, mode=R.REF_COMPAT, Nmc=Nmc, corrL=corrL,
                          xRange=(10, 10))
        kU, phU = R.reference_draws(p, R.reference_seed(11), calc)
        po = O.make_plan(2, O.RANDOMIZATION, O.REF_COMPAT, [10, 5])
        k = O.kpoints(po, kU)
        assert np.array_equal(kP, k)
        ref = O.shinozuka(po, x / corrL, k, sk, phU, calc=calc)
        tol = 1e-12
    else:
        p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, Nmc=Nmc, corrL=(0.5, 0.5, 0.5),
                          xRange=(10, 10, 10))
        _, U = R.reference_draws(p, R.reference_seed(7))
        po = O.make_plan(3, O.ISOTROPIC, O.FIXED, [20, 20, 20])
        k = O.kpoints(po)
        assert np.array_equal(kP, k)
        ref = O.isotropic(po, x / 0.5, k, sk, U)
        tol = 1e-12
    assert report(f"fortran driver case {case}", field, ref) <= tol


# ------------------------------------------------------------ full-size cfg3
def test_cfg3_full_size_properties(ctx):
    """2,000,000 points x 256 realizations x 4096 wavevectors on the MFMA
    form with the device generator: sampled points match the FP64 oracle,
    per-realization moments match N(0,1) and the lag correlation the target."""
    import torch
    N, Nmc, nk = 2_000_000, 256, 4096
    corrL = np.array([2.0, 2.0, 0.5])
    x = W.gaussian_mixture_points(N, 11)
    p = R.make_params(nDim=3, method=R.RANDOMIZATION, mode=R.FIXED, form=R.FORM_MFMA_F16,
                      kSampling=R.KSAMP_DENSITY, nKDensity=nk, Nmc=Nmc, corrL=tuple(corrL),
                      xRange=(1, 1, 1), seed=11)
    plan = R.FieldPlan(ctx, p)
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((Nmc, N), dtype=torch.float32, device="cuda")
    plan.execute(dx, out)
    ctx.sync()
    # the same uniforms: device stream tags 1 (wavevectors) and 2 (phases),
    # whose equality with the host mirror is tested above
    info = plan.info
    du = torch.empty(info.nKUniform + Nmc * nk, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 11, 1, info.nKUniform, du)
    R.device_uniforms(ctx, 11, 2, Nmc * nk, du[info.nKUniform:])
    ctx.sync()
    hu = du.cpu().numpy()
    kU, phU = hu[: info.nKUniform], hu[info.nKUniform:].reshape(Nmc, nk)
    sel = np.random.default_rng(0).choice(N, 48, replace=False)
    po = O.make_plan(3, O.RANDOMIZATION, O.FIXED, [1, 1, 1], ksamp=O.KDENSITY, nKDensity=nk)
    k = O.kpoints(po, kU)
    ref = O.shinozuka(po, x[sel] / corrL, k, O.skvec(po, k), phU)
    got = out[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
    assert report("cfg3 full-size sampled points mfma-f16", got, ref) <= 1e-5
    avg, sd, g = R.stats(ctx, out, N, Nmc, R.DTYPE_F32)
    print(f"[parity] cfg3 full: global mean={g[0]:.4f} std={g[1]:.4f}", flush=True)
    assert abs(g[0]) < 0.05 and abs(g[1] - 1.0) < 0.05
    del out, dx
    plan.close()
