"""Full-size GPU checks of the localization and isotropic-lognormal workloads,
through size-independent properties and sampled points against the FP64
oracle on the same device draws.

Localization: 512 x 512 x 256 grid at spacing 0.25, 4 x 4 x 2 uneven
subdomains sharing 7-point overlaps, Shinozuka FIXED per subdomain on its own
lattice (seed 100 + field number, device stream), Nmc 16, FP32 grid form in
grouped launches, variance-preserving blend (calls_RF.f90:232-369,
localization_RF.f90:199-263).
  - sampled points of four subdomains (two corner ones) vs the oracle: rel RMS
    <= 1e-5
  - blended values at overlap points (two, four and eight covering
    subdomains) vs the oracle's fields weighted by the oracle's partition of
    unity: rel RMS <= 1e-5
  - the variance over a 7-point overlap slab and over an interior slab both
    equal the lattice variance ampMult^2/2 sum S within 3 % (Monte-Carlo)

Isotropic + lognormal: 128^3 grid at spacing 0.1, corrL 0.5, 286 shells,
Nmc 64, avg 1, var 0.25 (randomFieldND.f90:139-178,368-514):
  - FP32 direct (per-realization directions) and MFMA BF16 hi/lo (shared
    directions) with the fused lognormal map: sampled points vs the oracle,
    rel RMS <= 1e-5 resp. 4e-5
  - the field's mean and variance equal the lognormal closed forms for the
    lattice's Gaussian variance within Monte-Carlo error.
"""
import math

import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _sub_draws(ctx, s, kN, Nmc):
    import torch
    du = torch.empty(Nmc * kN, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 100 + s + 1, 2, Nmc * kN, du)
    ctx.sync()
    return du.cpu().numpy().reshape(Nmc, kN)


def test_cfg4_full_size(ctx):
    import torch
    groupN, parts, nOvl, step, Nmc = (512, 512, 256), (4, 4, 2), 7, (0.25, 0.25, 0.25), 16
    origins, sizes, masks = W.cfg4_layout(groupN, parts, nOvl)
    plans, fields, geo = [], [], []
    for s in range(len(sizes)):
        n = [int(v) for v in sizes[s]]
        o = [origins[s][a] * step[a] for a in range(3)]
        ext = [(n[a] - 1) * step[a] for a in range(3)]
        plans.append(R.FieldPlan(ctx, R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32,
                                                    Nmc=Nmc, xRange=ext, seed=100 + s + 1)))
        fields.append(torch.empty((Nmc, int(np.prod(n))), dtype=torch.float32, device="cuda"))
        geo.append((o, n, ext))
    g = torch.empty((Nmc, int(np.prod(groupN))), dtype=torch.float32, device="cuda")
    batch = R.GridBatch(ctx, plans, [x[0] for x in geo], [step] * len(geo), [x[1] for x in geo], fields,
                        R.DTYPE_F32)
    bplan = R.BlendPlan(ctx, groupN, Nmc, origins, sizes, [nOvl] * 3, masks, R.FIXED, fields, R.DTYPE_F32)
    batch.execute()
    bplan.execute(g)
    ctx.sync()
    print(f"[fullsize] cfg4: N_k of subdomain 0 = {plans[0].info.kNTotal}", flush=True)

    cache = {}

    def oracle_at(s, gidx):
        """FP64 oracle of subdomain s at group indices gidx (m, 3)."""
        if s not in cache:
            o, n, ext = geo[s]
            po = O.make_plan(3, O.SHINOZUKA, O.FIXED, ext)
            k = O.kpoints(po)
            cache[s] = (po, k, O.skvec(po, k), _sub_draws(ctx, s, po.kNTotal, Nmc))
        po, k, sk, phU = cache[s]
        x = np.asarray(gidx, np.float64) * np.array(step)
        return O.shinozuka(po, x, k, sk, phU, threads=16)

    # sampled points of subdomains 0 and 31 (corners) and two inner ones
    rng = np.random.default_rng(4)
    for s in (0, 9, 22, 31):
        n = sizes[s]
        loc = np.stack([rng.integers(0, n[a], 12) for a in range(3)], axis=1)
        loc[0] = [0, 0, 0]
        loc[1] = n - 1
        flat = loc[:, 0] + n[0] * (loc[:, 1] + n[1] * loc[:, 2])
        got = fields[s][:, torch.from_numpy(flat).cuda()].cpu().numpy().astype(np.float64)
        ref = oracle_at(s, loc + origins[s])
        e = rel_rms(got, ref)
        print(f"[fullsize] cfg4 subdomain {s}: rel_rms={e:.3e}", flush=True)
        assert e <= 1e-5

    # blended values where 2, 4 and 8 subdomains overlap, and inside one
    pts = np.array([[128, 60, 60], [60, 128, 60], [60, 60, 128], [128, 128, 60], [128, 128, 128],
                    [383, 255, 130], [126, 131, 127], [300, 20, 200], [3, 3, 3], [511, 511, 255]])
    pus = {}
    exp_vals, got_vals = [], []
    gflat = pts[:, 0] + groupN[0] * (pts[:, 1] + groupN[1] * pts[:, 2])
    gv = g[:, torch.from_numpy(gflat).cuda()].cpu().numpy().astype(np.float64)
    for i, p in enumerate(pts):
        cover = [s for s in range(len(sizes)) if np.all(origins[s] <= p) and np.all(p < origins[s] + sizes[s])]
        acc = np.zeros(Nmc)
        for s in cover:
            if s not in pus:
                pus[s] = O.unity_partition(sizes[s], [nOvl] * 3, int(masks[s]), O.FIXED)
            l = p - origins[s]
            w = pus[s][l[0] + sizes[s][0] * (l[1] + sizes[s][1] * l[2])]
            acc += w * oracle_at(s, p[None, :])[:, 0]
        print(f"[fullsize] cfg4 blend point {p.tolist()}: {len(cover)} subdomains", flush=True)
        exp_vals.append(acc)
        got_vals.append(gv[:, i])
    assert sorted({len([s for s in range(len(sizes)) if np.all(origins[s] <= p) and
                        np.all(p < origins[s] + sizes[s])]) for p in pts}) == [1, 2, 4, 8]
    e = rel_rms(np.array(got_vals), np.array(exp_vals))
    print(f"[fullsize] cfg4 blended points: rel_rms={e:.3e}", flush=True)
    assert e <= 1e-5

    # variance preserved across the overlaps (FIXED weights)
    po = cache[0][0]
    target = po.ampMult ** 2 / 2 * cache[0][2].sum()
    g4 = g.view(Nmc, groupN[2], groupN[1], groupN[0])
    v_ovl = float(g4[:, :, :, 125:132].double().var())
    v_in = float(g4[:, :, :, 40:80].double().var())
    v_z = float(g4[:, 125:132, :, :].double().var())
    print(f"[fullsize] cfg4 variance target={target:.4f} x-overlap={v_ovl:.4f} z-overlap={v_z:.4f} "
          f"interior={v_in:.4f}", flush=True)
    for v in (v_ovl, v_in, v_z):
        assert abs(v / target - 1.0) < 0.03
    del g4
    bplan.close()
    batch.close()
    for pl in plans:
        pl.close()


@pytest.mark.parametrize("form", [R.FORM_FP32, R.FORM_MFMA_BF16])
def test_cfg2_full_size(ctx, form):
    import torch
    n, Nmc, corrL = 128, 64, 0.5
    x = W.grid_points((n, n, n), (0.1, 0.1, 0.1), (0.0, 0.0, 0.0))
    shared = 1 if form != R.FORM_FP32 else 0
    p = R.make_params(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, form=form, Nmc=Nmc, corrL=(corrL,) * 3,
                      xRange=(12.7,) * 3, isoShared=shared, margiFirst=R.FOM_LOGNORMAL, fieldAvg=1.0,
                      fieldVar=0.25, seed=7)
    plan = R.FieldPlan(ctx, p)
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((Nmc, x.shape[0]), dtype=torch.float32, device="cuda")
    plan.execute(dx, out)
    ctx.sync()
    info = plan.info
    kN, nv = info.kNTotal, info.nPhaseUniform
    du = torch.empty(Nmc * nv, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 7, 2, Nmc * nv, du)
    ctx.sync()
    U = du.cpu().numpy().reshape(Nmc, nv)
    if shared:
        U[:, : 2 * kN] = U[0, : 2 * kN]          # phi and theta of realization 1 for all
    po = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    kr = O.kpoints(po)
    sk = O.skvec(po, kr)
    assert po.kNTotal == kN == 286
    sel = np.random.default_rng(2).choice(x.shape[0], 40, replace=False)
    ref = O.marginal(2, 1.0, 0.25, O.isotropic(po, x[sel] / corrL, kr, sk, U, threads=16))
    got = out[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
    e = rel_rms(got, ref)
    tol = 4e-5 if form == R.FORM_MFMA_BF16 else 1e-5
    print(f"[fullsize] cfg2 form={form}: rel_rms={e:.3e}", flush=True)
    assert e <= tol
    # lognormal moments: the Gaussian field's variance given the direction
    # draws is ampMult^2 sum_j S_j k_j^2 sin(phi_j) E[gamma^2] / 2 (shared
    # directions), or its expectation ampMult^2 sum S k^2 / pi
    k = kr[:, 0]
    if shared:
        v = po.ampMult ** 2 * np.sum(sk * k ** 2 * np.sin(math.pi * U[0, :kN])) / 2
    else:
        v = po.ampMult ** 2 * np.sum(sk * k ** 2) / math.pi
    s2 = math.log(1.0 + 0.25)
    mu = -s2 / 2
    mean_exp = math.exp(mu + s2 * v / 2)
    var_exp = mean_exp ** 2 * (math.exp(s2 * v) - 1.0)
    avg, sd, gl = R.stats(ctx, out, x.shape[0], Nmc, R.DTYPE_F32)
    print(f"[fullsize] cfg2 form={form}: mean={gl[0]:.4f} (closed form {mean_exp:.4f}) "
          f"var={gl[1] ** 2:.4f} (closed form {var_exp:.4f})", flush=True)
    assert abs(gl[0] - mean_exp) < 0.01
    assert abs(gl[1] ** 2 / var_exp - 1.0) < 0.06
    del out, dx
    plan.close()
