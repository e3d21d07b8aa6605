"""GPU parity of the structured-grid form (axis-wise contraction of the
Shinozuka sum) and of the localization pipeline built on it: subdomains
generated on their own grids and lattices, then blended.  The oracle sums the
same wavevectors point by point in FP64 (randomFieldND.f90:276-330 order)."""
from pathlib import Path

import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd import workloads as W

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def grid_case(ctx, nstep, step, origin, corrL, Nmc, seed, dtype=np.float32, margi=R.FOM_NONE):
    import torch
    d = len(nstep)
    ext = [(nstep[a] - 1) * step[a] for a in range(d)]
    p = R.make_params(nDim=d, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=Nmc,
                      corrL=corrL, xRange=ext, margiFirst=margi, fieldAvg=1.0, fieldVar=0.25)
    _, phU = R.reference_draws(p, R.reference_seed(seed))
    plan = R.FieldPlan(ctx, p, None, phU)
    out = torch.empty((Nmc, int(np.prod(nstep))), dtype=torch.float32 if dtype == np.float32 else torch.float64,
                      device="cuda")
    plan.execute_grid(origin, step, nstep, out)
    ctx.sync()
    x = W.grid_points(nstep, step, origin)
    po = O.make_plan(d, O.SHINOZUKA, O.FIXED, [ext[a] / corrL[a] for a in range(d)])
    k = O.kpoints(po)
    ref = O.shinozuka(po, x / np.array(corrL), k, O.skvec(po, k), phU)
    if margi != R.FOM_NONE:
        ref = O.marginal(margi, 1.0, 0.25, ref)
    got = out.cpu().numpy().astype(np.float64)
    plan.close()
    return got, ref


@pytest.mark.parametrize("nstep,step,corrL", [
    ((20, 18, 16), (0.25, 0.25, 0.25), (1.0, 1.2, 0.8)),
    ((33, 7, 40), (0.3, 0.5, 0.2), (1.0, 1.0, 1.0)),
    ((41, 29), (0.2, 0.3), (0.7, 1.3)),
    # x1 = 129: the middle contraction's 128-column tiles straddle (j3, x1)
    # block boundaries, the last contraction's end in a 5-column tile
    ((129, 5, 3), (0.2, 0.4, 0.5), (1.0, 1.0, 1.0)),
    # x2 = 140: partial 64-row tiles of the middle contraction
    ((7, 140, 9), (0.5, 0.1, 0.4), (1.0, 1.0, 1.0)),
    # x1 of 129..160 runs the first contraction on 160-column tiles: a full
    # tile, a short last column group, and the 2-D form
    ((160, 6, 4), (0.1, 0.4, 0.5), (1.0, 1.0, 1.0)),
    ((137, 70, 5), (0.15, 0.2, 0.5), (1.1, 0.9, 1.0)),
    ((150, 23), (0.1, 0.3), (1.0, 1.2)),
    # a last axis of 129..160 points: the last contraction's short last row tile
    ((16, 8, 150), (0.3, 0.4, 0.1), (1.0, 1.0, 1.0)),
    ((24, 9, 160), (0.25, 0.3, 0.1), (0.9, 1.0, 1.1)),
    ((100, 140), (0.1, 0.1), (1.0, 0.9)),
])
def test_grid_form_matches_oracle(ctx, nstep, step, corrL):
    d = len(nstep)
    got, ref = grid_case(ctx, nstep, step, (1.5, -2.0, 0.25)[:d], corrL[:d], 3, 5)
    e = rel_rms(got, ref)
    print(f"[parity] grid {nstep}: rel_rms={e:.3e} max_abs={np.max(np.abs(got - ref)):.3e}", flush=True)
    assert e <= 1e-5


def test_grid_form_lognormal_and_fp64_out(ctx):
    got, ref = grid_case(ctx, (12, 11, 10), (0.25, 0.25, 0.25), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 2, 9,
                         dtype=np.float64, margi=R.FOM_LOGNORMAL)
    assert rel_rms(got, ref) <= 1e-5


def test_grid_form_rejects_other_methods(ctx):
    import torch
    for method, mode in ((R.SHINOZUKA, R.REF_COMPAT), (R.RANDOMIZATION, R.FIXED)):
        p = R.make_params(nDim=3, method=method, mode=mode, Nmc=1, xRange=(2, 2, 2))
        plan = R.FieldPlan(ctx, p)
        out = torch.empty((1, 27), dtype=torch.float32, device="cuda")
        with pytest.raises(R.RFError) as e:
            plan.execute_grid((0, 0, 0), (1, 1, 1), (3, 3, 3), out)
        assert e.value.code == R.RF_E_UNSUPPORTED
        plan.close()


def test_localization_pipeline_small(ctx):
    """Uneven tensor-product subdomains (seed 100 + field index, last axis
    fastest) generated on the grid form and blended with variance-preserving
    weights, against the oracle doing the same per point."""
    import torch
    groupN = (40, 36, 20)
    parts = (3, 2, 2)
    nOvl = 5
    step = (0.25, 0.25, 0.25)
    Nmc = 2
    origins, sizes, masks = W.cfg4_layout(groupN, parts, nOvl)
    fields_d, fields_h = [], []
    for s in range(len(sizes)):
        n = [int(v) for v in sizes[s]]
        o = [origins[s][a] * step[a] for a in range(3)]
        ext = [(n[a] - 1) * step[a] for a in range(3)]
        p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=Nmc,
                          xRange=ext, seed=100 + s + 1)
        plan = R.FieldPlan(ctx, p)
        f = torch.empty((Nmc, int(np.prod(n))), dtype=torch.float32, device="cuda")
        plan.execute_grid(o, step, n, f)
        ctx.sync()
        # the oracle, same device uniforms (tag 2 stream of the subdomain seed)
        info = plan.info
        du = torch.empty(Nmc * info.kNTotal, dtype=torch.float64, device="cuda")
        R.device_uniforms(ctx, 100 + s + 1, 2, Nmc * info.kNTotal, du)
        ctx.sync()
        phU = du.cpu().numpy().reshape(Nmc, info.kNTotal)
        po = O.make_plan(3, O.SHINOZUKA, O.FIXED, ext)
        k = O.kpoints(po)
        x = W.grid_points(n, step, o)
        ref = O.shinozuka(po, x, k, O.skvec(po, k), phU)
        assert rel_rms(f.cpu().numpy().astype(np.float64), ref) <= 1e-5
        fields_d.append(f)
        fields_h.append(ref)
        plan.close()
    g = torch.empty((Nmc, int(np.prod(groupN))), dtype=torch.float32, device="cuda")
    R.blend(ctx, groupN, Nmc, origins, sizes, [nOvl] * 3, masks, R.FIXED, fields_d, g, R.DTYPE_F32)
    ctx.sync()
    gref = O.blend(groupN, Nmc, origins, sizes, [nOvl] * 3, masks, O.FIXED, fields_h)
    e = rel_rms(g.cpu().numpy().astype(np.float64), gref)
    print(f"[parity] localization pipeline: rel_rms={e:.3e}", flush=True)
    assert e <= 1e-5


def _grid_fields(ctx, specs, dtype):
    import torch
    plans, outs, geo = [], [], []
    for (nstep, step, origin, corrL, Nmc, seed) in specs:
        d = len(nstep)
        ext = [(nstep[a] - 1) * step[a] for a in range(d)]
        p = R.make_params(nDim=d, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=Nmc, corrL=corrL,
                          xRange=ext, seed=seed, margiFirst=R.FOM_GAUSSIAN, fieldAvg=0.5, fieldVar=2.0)
        plans.append(R.FieldPlan(ctx, p))
        outs.append(torch.empty((Nmc, int(np.prod(nstep))), dtype=dtype, device="cuda"))
        geo.append((origin, step, nstep))
    return plans, outs, geo


SPECS = [((20, 18, 16), (0.25, 0.25, 0.25), (1.5, -2.0, 0.25), (1.0, 1.2, 0.8), 3, 11),
         ((33, 7), (0.3, 0.5), (0.0, 1.0), (1.0, 1.0), 5, 12),
         ((9, 40, 12), (0.5, 0.2, 0.3), (0.0, 0.0, 0.0), (0.7, 1.3, 1.0), 1, 13),
         ((64, 64, 8), (0.25, 0.25, 0.25), (4.0, 0.0, 1.0), (1.0, 1.0, 1.0), 16, 14)]


def test_grid_batch_equals_single(ctx):
    """All fields in three grouped launches give the same bits as one
    rf_plan_execute_grid per field."""
    import torch
    plans, outs, geo = _grid_fields(ctx, SPECS, torch.float32)
    singles = []
    for pl, o, (org, st, ns) in zip(plans, outs, geo):
        pl.execute_grid(org, st, ns, o)
        ctx.sync()
        singles.append(o.cpu().numpy().copy())
        o.zero_()
    b = R.GridBatch(ctx, plans, [g[0] for g in geo], [g[1] for g in geo], [g[2] for g in geo], outs, R.DTYPE_F32)
    b.execute()
    ctx.sync()
    for i, o in enumerate(outs):
        assert np.array_equal(o.cpu().numpy(), singles[i]), i
    b.close()
    for pl in plans:
        pl.close()


def test_graph_replay_of_batch_and_blend(ctx):
    """Generation of every subdomain (grouped launches) and the blend,
    captured once and replayed; each replay, with eager work and
    synchronisation in between, matches the eager result."""
    import torch
    groupN, parts, nOvl, step, Nmc = (40, 36, 20), (3, 2, 2), 5, (0.25, 0.25, 0.25), 2
    origins, sizes, masks = W.cfg4_layout(groupN, parts, nOvl)
    specs = []
    for s in range(len(sizes)):
        n = tuple(int(v) for v in sizes[s])
        specs.append((n, step, tuple(origins[s][a] * step[a] for a in range(3)), (1.0, 1.0, 1.0), Nmc, 100 + s + 1))
    plans, outs, geo = _grid_fields(ctx, specs, torch.float32)
    g = torch.empty((Nmc, int(np.prod(groupN))), dtype=torch.float32, device="cuda")
    batch = R.GridBatch(ctx, plans, [x[0] for x in geo], [x[1] for x in geo], [x[2] for x in geo], outs, R.DTYPE_F32)
    bp = R.BlendPlan(ctx, groupN, Nmc, origins, sizes, [nOvl] * 3, masks, R.FIXED, outs, R.DTYPE_F32)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    batch.execute(sp)
    bp.execute(g, sp)
    torch.cuda.synchronize()
    eager = g.cpu().numpy().copy()
    eager_f = [o.cpu().numpy().copy() for o in outs]
    # the one-shot blend gives the same group field
    g2 = torch.empty_like(g)
    R.blend(ctx, groupN, Nmc, origins, sizes, [nOvl] * 3, masks, R.FIXED, outs, g2, R.DTYPE_F32)
    ctx.sync()
    assert np.array_equal(g2.cpu().numpy(), eager)
    graph = R.Graph(ctx, sp)
    with graph:
        batch.execute(sp)
        bp.execute(g, sp)
    for rep in range(3):
        g.fill_(float(rep + 7))
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        graph.launch(sp)
        torch.cuda.synchronize()
        assert np.array_equal(g.cpu().numpy(), eager), rep
        for i, o in enumerate(outs):
            assert np.array_equal(o.cpu().numpy(), eager_f[i]), (rep, i)
    graph.close()
    bp.close()
    batch.close()
    for pl in plans:
        pl.close()


def test_generated_grid_to_hdf5(ctx, tmp_path):
    """A generated grid's realization 1 through the HDF5 writer, read back by
    h5dump, equals the device field (x fastest, HDF5 dims last axis first)."""
    import re
    import shutil
    import subprocess
    import torch
    from randomfield_amd import h5io
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not h5io.LIB.exists() or not Path(h5dump).exists():
        pytest.skip("HDF5 C library or h5dump absent")
    n, step = (20, 12, 9), (0.25, 0.25, 0.25)
    p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=2, xRange=(5, 3, 2), seed=4)
    plan = R.FieldPlan(ctx, p)
    out = torch.empty((2, int(np.prod(n))), dtype=torch.float64, device="cuda")
    plan.execute_grid((0.0, 0.0, 0.0), step, n, out)
    ctx.sync()
    plan.close()
    h5, _ = h5io.write_simple(tmp_path, "gen", out[0], n, step, (0.0, 0.0, 0.0), n)
    txt = subprocess.run([h5dump, "-d", "/samples", "-y", "-w", "0", "-m", "%.17g", h5], capture_output=True,
                         text=True, check=True).stdout
    body = txt.split("DATA {", 1)[1].rsplit("}", 3)[0]
    vals = np.array([float(v) for v in re.split(r"[,\s]+", body.strip()) if v])
    assert np.array_equal(vals, out[0].cpu().numpy())


@pytest.mark.parametrize("method", ["shinozuka", "fft"])
def test_repeated_grid_calls_reuse_coefficients(ctx, method):
    """A plan keeps its folded coefficients between grid executions: a second
    grid (other origin and step; for method 4 another step, hence another
    scale) gives the bits of a fresh plan."""
    import torch

    def make():
        if method == "fft":
            p = R.make_params(nDim=3, method=R.FFT, mode=R.FIXED, form=R.FORM_FP32, Nmc=2, xRange=(3.0, 2.0, 1.5),
                              gridN=(13, 9, 7), seed=8)
        else:
            p = R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED, form=R.FORM_FP32, Nmc=2, xRange=(4, 3, 2),
                              seed=8)
        return R.FieldPlan(ctx, p)
    grids = [((0.0, 0.0, 0.0), (0.25, 0.25, 0.25), (13, 9, 7)), ((0.5, -1.0, 2.0), (0.3, 0.2, 0.4), (13, 9, 7))]
    plan = make()
    outs = []
    for org, st, ns in grids:
        o = torch.empty((2, int(np.prod(ns))), dtype=torch.float32, device="cuda")
        plan.execute_grid(org, st, ns, o)
        ctx.sync()
        outs.append(o.cpu().numpy())
    plan.close()
    for (org, st, ns), got in zip(grids, outs):
        fresh = make()
        o = torch.empty((2, int(np.prod(ns))), dtype=torch.float32, device="cuda")
        fresh.execute_grid(org, st, ns, o)
        ctx.sync()
        fresh.close()
        assert np.array_equal(o.cpu().numpy(), got)


REDRAW_SPECS = {
    "fold": SPECS[:2] + [((12, 10, 9), (0.3, 0.3, 0.3), (1.0, 0.0, -1.0), (1.0, 1.0, 1.0), 4, 21)],
    # first axes of 129..160 points: every first contraction on 64x160 tiles,
    # which then form the fresh coefficients themselves (no fold)
    "formed": [((131, 6, 5), (0.1, 0.3, 0.4), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 3, 41),
               ((140, 7), (0.08, 0.3), (0.5, 0.0), (1.0, 1.2), 2, 42),
               ((135, 5, 4), (0.1, 0.3, 0.3), (1.0, 0.0, -1.0), (1.0, 1.0, 1.0), 4, 43)],
}


@pytest.mark.parametrize("which", ["fold", "formed"])
def test_grid_batch_redraw(ctx, which):
    """Fresh realizations in a batch: redraw(0) gives the plans' own fields
    bit for bit; redraw(2) gives realizations 2 Nmc .. 3 Nmc - 1 of each
    plan's device stream (checked against the oracle on those draws); a graph
    captured with the draw counter draws the next set on every replay, equal
    to the eager redraw of that index.  Both ways of forming the fresh
    coefficients: the grouped fold, and the first contraction itself."""
    import torch
    specs = REDRAW_SPECS[which]
    plans, outs, geo = _grid_fields(ctx, specs, torch.float32)
    b = R.GridBatch(ctx, plans, [g[0] for g in geo], [g[1] for g in geo], [g[2] for g in geo], outs, R.DTYPE_F32)
    b.execute()
    ctx.sync()
    own = [o.cpu().numpy().copy() for o in outs]
    b.redraw(0)
    b.execute()
    ctx.sync()
    for i, o in enumerate(outs):
        assert np.array_equal(o.cpu().numpy(), own[i]), i
    eager = {}
    for d in (1, 2, 3):
        b.redraw(d)
        b.execute()
        ctx.sync()
        eager[d] = [o.cpu().numpy().copy() for o in outs]
    assert not np.array_equal(eager[2][0], own[0])
    # draw 2 of field 2 against the oracle on the same stream elements
    (nstep, step, origin, corrL, Nmc, seed) = specs[2]
    ext = [(nstep[a] - 1) * step[a] for a in range(3)]
    kN = plans[2].info.kNTotal
    du = torch.empty(3 * Nmc * kN, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, seed, 2, 3 * Nmc * kN, du)
    ctx.sync()
    phU = du.cpu().numpy()[2 * Nmc * kN:].reshape(Nmc, kN)
    po = O.make_plan(3, O.SHINOZUKA, O.FIXED, [ext[a] / corrL[a] for a in range(3)])
    k = O.kpoints(po)
    x = W.grid_points(nstep, step, origin)
    ref = O.shinozuka(po, x / np.array(corrL), k, O.skvec(po, k), phU)
    ref = O.marginal(1, 0.5, 2.0, ref)          # the specs' gaussian marginal (avg 0.5, var 2)
    e = rel_rms(eager[2][2].astype(np.float64), ref)
    print(f"[parity] batch redraw 2: rel_rms={e:.3e}", flush=True)
    assert e <= 1e-5
    # graph: counter-driven draws
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    b.redraw(1, sp)                      # sets the counter to 1 (and draws set 1)
    torch.cuda.synchronize()
    graph = R.Graph(ctx, sp)
    with graph:
        b.redraw(R.DRAW_NEXT, sp)
        b.execute(sp)
    # capturing launched nothing: the counter is still 1
    for rep, d in enumerate((1, 2, 3)):
        for o in outs:
            o.fill_(float(rep + 5))
        torch.cuda.synchronize()
        graph.launch(sp)
        torch.cuda.synchronize()
        for i, o in enumerate(outs):
            assert np.array_equal(o.cpu().numpy(), eager[d][i]), (rep, i)
    graph.close()
    # the autotuner after a redraw keeps the current realizations, whichever
    # tile kind it picks
    b.redraw(2)
    b.autotune()
    for o in outs:
        o.fill_(-1.0)
    b.execute()
    ctx.sync()
    for i, o in enumerate(outs):
        assert np.array_equal(o.cpu().numpy(), eager[2][i]), i
    b.close()
    for pl in plans:
        pl.close()


@pytest.mark.parametrize("groupN,parts,nOvl", [
    ((40, 36, 20), (3, 2, 2), 5),
    # subdomains of 131/132 points along x3 (a short last row tile in the last
    # contraction) and along x1 (64x160 tiles in the first)
    ((40, 36, 256), (2, 2, 2), 7),
    ((256, 20, 16), (2, 1, 2), 7),
])
def test_grid_batch_autotune_and_fused_blend(ctx, groupN, parts, nOvl):
    """The tile autotuner and the blend fused into the last contraction leave
    every group value and every overlap value of the fields unchanged, bit for
    bit; a fused graph with fresh draws matches the eager fused result; the
    blend plan refuses another group buffer while fused; unfusing restores
    the separate blend."""
    import torch
    step, Nmc = (0.25, 0.25, 0.25), 3
    origins, sizes, masks = W.cfg4_layout(groupN, parts, nOvl)
    specs = []
    for s in range(len(sizes)):
        n = tuple(int(v) for v in sizes[s])
        specs.append((n, step, tuple(origins[s][a] * step[a] for a in range(3)), (1.0, 1.0, 1.0), Nmc, 200 + s))
    plans, outs, geo = _grid_fields(ctx, specs, torch.float32)
    batch = R.GridBatch(ctx, plans, [x[0] for x in geo], [x[1] for x in geo], [x[2] for x in geo], outs, R.DTYPE_F32)
    bp = R.BlendPlan(ctx, groupN, Nmc, origins, sizes, [nOvl] * 3, masks, R.FIXED, outs, R.DTYPE_F32)
    g = torch.empty((Nmc, int(np.prod(groupN))), dtype=torch.float32, device="cuda")
    batch.execute()
    bp.execute(g)
    ctx.sync()
    ref_g = g.cpu().numpy().copy()
    ref_f = [o.cpu().numpy().copy() for o in outs]
    chosen = batch.autotune()
    print(f"[grid] autotune chose tile kinds {chosen}", flush=True)
    assert len(chosen) == 3 and all(c in (1, 2, 3, 4, 5) for c in chosen)
    g.fill_(-7.0)
    batch.execute()
    bp.execute(g)
    ctx.sync()
    assert np.array_equal(g.cpu().numpy(), ref_g)
    # fused: interior points never reach the field buffers
    g2 = torch.full_like(g, 1234.5)
    for o in outs:
        o.fill_(9876.5)
    batch.fuse_blend(bp, g2)
    batch.execute()
    bp.execute(g2)
    ctx.sync()
    got = g2.cpu().numpy()
    print(f"[grid] fused blend: max |diff| {np.max(np.abs(got - ref_g)):.3e}", flush=True)
    assert np.array_equal(got, ref_g)
    n_interior = 0
    for i, o in enumerate(outs):
        f = o.cpu().numpy()
        written = f != np.float32(9876.5)
        assert np.array_equal(f[written], ref_f[i][written])
        n_interior += int((~written).sum())
    assert n_interior > 0
    with pytest.raises(R.RFError):
        bp.execute(g)
    # fused graph with fresh draws against eager fused redraws
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    eager = {}
    for d in (1, 2):
        batch.redraw(d, sp)
        batch.execute(sp)
        bp.execute(g2, sp)
        torch.cuda.synchronize()
        eager[d] = g2.cpu().numpy().copy()
    batch.redraw(1, sp)
    torch.cuda.synchronize()
    graph = R.Graph(ctx, sp)
    with graph:
        batch.redraw(R.DRAW_NEXT, sp)
        batch.execute(sp)
        bp.execute(g2, sp)
    for d in (1, 2):
        g2.fill_(float(d))
        torch.cuda.synchronize()
        graph.launch(sp)
        torch.cuda.synchronize()
        assert np.array_equal(g2.cpu().numpy(), eager[d]), d
    graph.close()
    # unfused again: the separate blend on any group buffer
    batch.fuse_blend(bp, None)
    batch.redraw(0)
    batch.execute()
    g.fill_(3.0)
    bp.execute(g)
    ctx.sync()
    assert np.array_equal(g.cpu().numpy(), ref_g)
    bp.close()
    batch.close()
    for pl in plans:
        pl.close()
