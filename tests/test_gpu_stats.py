"""GPU statistics: the directional correlation and corrL estimate of
statistics_RF against the oracle, and generated fields against the target
correlation R(tau) = exp(-pi tau^2 / L^2) within Monte-Carlo error."""
import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nstep", [(40,), (24, 17), (20, 13, 9)])
def test_corr_device_matches_oracle(ctx, dtype, nstep):
    import torch
    rng = np.random.default_rng(sum(nstep))
    N = int(np.prod(nstep))
    f = rng.normal(0.4, 2.0, (3, N)).astype(dtype)
    d = torch.from_numpy(f).cuda()
    xs = [0.1 * (a + 1) for a in range(len(nstep))]
    for r in (0, 2):
        u = f[r].astype(np.float64)
        avg, sd = float(u.mean()), float(u.std())
        got, cl = R.corr(ctx, d, R.DTYPE_F64 if dtype == np.float64 else R.DTYPE_F32, nstep, xs, avg, sd, r)
        ref = O.corr(u, nstep, avg, sd)
        rcl = O.corrL(ref, nstep, xs)
        for a in range(len(nstep)):
            assert np.allclose(got[a], ref[a], rtol=1e-12, atol=1e-13), (r, a)
        assert np.allclose(cl, rcl, rtol=1e-12, atol=1e-13)


def test_generated_field_moments_and_correlation(ctx):
    """Shinozuka fields (FIXED mode, grid form) on 128 x 128 x 64 points at
    spacing 0.1 L against the statistics of the discrete spectral sum they
    realise (the lattice, not the continuum: the variance is ampMult^2/2 sum S
    with the FIXED trapezoid weights on the k = 0 planes, close to 1): mean within 4 standard
    errors of the box average, variance within 8 %, lag correlations of every
    axis within 0.06 of sum S cos(k tau) / sum S (times (n - m)/n, the bias
    of the reference's circular estimator), and the reference's corrL
    estimate within 10 % of the same estimator applied to that target."""
    import torch
    nstep, step, Nmc = (128, 128, 64), (0.1, 0.1, 0.1), 4
    ext = [(n - 1) * s for n, s in zip(nstep, step)]
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=Nmc, xRange=ext, seed=2024)
    plan = R.FieldPlan(ctx, p)
    info = plan.info
    N = int(np.prod(nstep))
    out = torch.empty((Nmc, N), dtype=torch.float32, device="cuda")
    plan.execute_grid((0.0, 0.0, 0.0), step, nstep, out)
    ctx.sync()
    plan.close()
    # statistics of the discrete sum on this lattice
    po = O.make_plan(3, O.SHINOZUKA, O.FIXED, ext)
    kt = O.kpoints(po)                                 # (N_k, 3)
    S = O.skvec(po, kt)
    k = kt.T
    w = po.ampMult ** 2 / 2.0 * S                     # variance of each term
    var_t = float(w.sum())
    box = np.ones(k.shape[1])
    for a in range(3):
        x = step[a] * np.arange(nstep[a])
        box *= np.abs(np.exp(1j * np.outer(k[a], x)).mean(axis=1)) ** 2
    mean_sd = float(np.sqrt((w * box).sum()))          # sd of one realization's box mean
    avg, sd, g = R.stats(ctx, out, N, Nmc, R.DTYPE_F32)
    print(f"[stats] mean {g[0]:.4f} (sd {mean_sd / np.sqrt(Nmc):.4f}) var {np.mean(sd ** 2):.4f} "
          f"(target {var_t - mean_sd ** 2:.4f})", flush=True)
    assert abs(g[0]) < 4 * mean_sd / np.sqrt(Nmc) + 0.01
    assert abs(np.mean(sd ** 2) / (var_t - mean_sd ** 2) - 1.0) < 0.08
    Rs, cls = [], []
    for r in range(Nmc):
        Rr, cl = R.corr(ctx, out, R.DTYPE_F32, nstep, step, float(avg[r]), float(sd[r]), r)
        Rs.append(Rr)
        cls.append(cl)
    cl = np.mean(cls, axis=0)
    for a in range(3):
        n = nstep[a]
        M = n // 2
        lags = np.arange(M) * step[a]
        rho = (w[None, :] * np.cos(np.outer(lags, k[a]))).sum(axis=1) / w.sum() * (n - np.arange(M)) / n
        cl_t = O.corrL([rho], [n], [step[a]])[0]
        Ra = np.mean([Rr[a] for Rr in Rs], axis=0)
        print(f"[stats] axis {a}: corrL estimate {cl[a]:.4f} (target {cl_t:.4f})", flush=True)
        assert abs(cl[a] / cl_t - 1.0) < 0.10
        for m in (0, 3, 5, 8, 12):
            assert abs(Ra[m] - rho[m]) < 0.06, (a, m, Ra[m], rho[m])


@pytest.mark.parametrize("nDim", [2, 3])
@pytest.mark.parametrize("gdt,odt", [(np.float64, np.float64), (np.float32, np.float32)])
def test_interpolate_matches_oracle(ctx, nDim, gdt, odt):
    """interpolateToMesh on the device, every realization, against the
    oracle; FP64 bit for bit (same operation order)."""
    import torch
    rng = np.random.default_rng(nDim)
    nstep = (33, 20, 12)[:nDim]
    xMin = (-1.0, 2.0, 0.5)[:nDim]
    xStep = (0.25, 0.5, 0.75)[:nDim]
    Nmc = 6
    grid = rng.standard_normal((Nmc, int(np.prod(nstep)))).astype(gdt)
    hi = np.array([xMin[a] + xStep[a] * (nstep[a] - 1) for a in range(nDim)])
    x = rng.uniform(np.array(xMin), hi, (20000, nDim))
    x[0] = hi
    ref, bad = O.interp(grid.astype(np.float64), nstep, xMin, xStep, x)
    assert bad == 0
    tt = {np.float64: torch.float64, np.float32: torch.float32}
    out = torch.empty((Nmc, x.shape[0]), dtype=tt[odt], device="cuda")
    outside = R.interpolate(ctx, nstep, xMin, xStep, Nmc, torch.from_numpy(grid).cuda(), torch.from_numpy(x).cuda(),
                            out)
    assert not outside
    got = out.cpu().numpy()
    if odt == np.float64:
        assert np.array_equal(got, ref)
    else:
        assert np.array_equal(got, ref.astype(np.float32))
    xo = np.concatenate([x[:10], (np.array(xMin) - 1.0)[None, :]])
    out2 = torch.empty((Nmc, xo.shape[0]), dtype=tt[odt], device="cuda")
    assert R.interpolate(ctx, nstep, xMin, xStep, Nmc, torch.from_numpy(grid).cuda(), torch.from_numpy(xo).cuda(),
                         out2)
    assert np.isnan(out2.cpu().numpy()[:, -1]).all()


def test_interpolate_box_equals_full_grid(ctx):
    """The bounding-box form (the reference reads only the nodes' box) gives
    the bits of the whole-grid form for nodes inside the box, borders
    included; nodes outside the box are flagged."""
    import torch
    rng = np.random.default_rng(11)
    nstep, xMin, xStep, Nmc = (30, 22, 14), (-1.0, 2.0, 0.5), (0.25, 0.5, 0.75), 3
    grid = rng.standard_normal((Nmc, int(np.prod(nstep))))
    mn, mx = (4, 3, 2), (19, 15, 9)
    lo = np.array([xMin[a] + (mn[a] - 1) * xStep[a] for a in range(3)])
    hi = np.array([xMin[a] + (mx[a] - 1) * xStep[a] for a in range(3)])
    x = rng.uniform(lo, hi, (5000, 3))
    x[0], x[1] = hi, lo
    x[2, 0] = hi[0]
    G = grid.reshape(Nmc, nstep[2], nstep[1], nstep[0])
    box = np.ascontiguousarray(G[:, mn[2] - 1:mx[2], mn[1] - 1:mx[1], mn[0] - 1:mx[0]].reshape(Nmc, -1))
    d_x = torch.from_numpy(x).cuda()
    full = torch.empty((Nmc, x.shape[0]), dtype=torch.float64, device="cuda")
    part = torch.empty_like(full)
    assert not R.interpolate(ctx, nstep, xMin, xStep, Nmc, torch.from_numpy(grid).cuda(), d_x, full)
    assert not R.interpolate_box(ctx, mn, mx, xMin, xStep, Nmc, torch.from_numpy(box).cuda(), d_x, part)
    assert np.array_equal(full.cpu().numpy(), part.cpu().numpy())
    xo = torch.from_numpy(np.array([hi + 1.0])).cuda()
    o = torch.empty((Nmc, 1), dtype=torch.float64, device="cuda")
    assert R.interpolate_box(ctx, mn, mx, xMin, xStep, Nmc, torch.from_numpy(box).cuda(), xo, o)


def test_interpolate_to_mesh_pipeline(ctx, tmp_path):
    """Generation -> HDF5 samples + attributes -> UNV mesh -> interpolateToMesh
    on the nodes' bounding box, against the oracle's interpolation of the
    whole grid (FP64, bit for bit)."""
    import torch
    from randomfield_amd import h5io
    if not h5io.LIB.exists():
        pytest.skip("HDF5 C library absent")
    nstep, step, x0 = (24, 18, 10), (0.25, 0.25, 0.5), (0.0, -1.0, 1.0)
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=1, xRange=(6, 4.5, 4.5),
                      seed=9)
    plan = R.FieldPlan(ctx, p)
    field = torch.empty((1, int(np.prod(nstep))), dtype=torch.float64, device="cuda")
    plan.execute_grid(x0, step, nstep, field)
    ctx.sync()
    plan.close()
    f = field[0].cpu().numpy()
    h5, _ = h5io.write_simple(tmp_path, "gen", f, nstep, step, x0, nstep)
    xmax = [x0[a] + (nstep[a] - 1) * step[a] for a in range(3)]
    h5io.write_attributes(h5, nDim=3, Nmc=1, method=2, seedStart=9, corrMod=1, margiFirst=1, xMinGlob=x0,
                          xMaxGlob=xmax, xStep=step, corrL=(1, 1, 1))
    # a brick mesh inside the grid, written as a universal file
    rng = np.random.default_rng(5)
    nodes = rng.uniform([0.6, -0.4, 1.7], [4.1, 2.3, 4.2], (700, 3))
    lines = ["    -1", "  2411"]
    for i, c in enumerate(nodes):
        lines += [f"{i + 1:10d}{1:10d}{1:10d}{11:10d}", "".join(f"{v:25.16E}".replace("E", "D") for v in c)]
    lines += ["    -1", "    -1", "  2412", f"{1:10d}{115:10d}{2:10d}{1:10d}{7:10d}{8:10d}",
              "".join(f"{v:10d}" for v in range(1, 9)), "    -1"]
    unv = tmp_path / "mesh.unv"
    unv.write_text("\n".join(lines) + "\n")
    coords, *_ = R.read_unv(unv, 3)
    got = h5io.interpolate_to_mesh(ctx, h5, coords)
    ref, bad = O.interp(f[None, :], nstep, x0, step, coords)
    assert bad == 0
    assert np.array_equal(got, ref[0])
