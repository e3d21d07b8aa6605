"""Benchmark of the spectral random-field hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): randomization
method, 2,000,000 unstructured points of an 8-component Gaussian mixture
(centres U[0,100]^3, std 5), corrL (2,2,0.5), N_k = 4096 wavevectors drawn
from the spectral density, Nmc = 256 realizations, FP16 hi/lo MFMA form with
FP32 accumulation.  A step is one rf_plan_execute over all points (points,
wavevectors and coefficient panels resident in HBM before timing starts).

Prints one JSON line: throughput in points x realizations / s, the MFMA
roofline of the kernel (HIP events on the launch stream), p50/p99 step
latency, the plain FP32 direct sum on the same workload, the error of sampled
points against the FP64 oracle, and the oracle's CPU throughput on a bounded
sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_F16_DENSE = 2.5e15      # MI355X dense FP16 MFMA, FLOP/s
N_POINTS, N_MC, N_K = 2_000_000, 256, 4096
CORR_L = (2.0, 2.0, 0.5)


def _threads() -> int:
    n = os.cpu_count() or 1
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return min(n, int(os.environ.get("OMP_NUM_THREADS", n)))


def _timed(fn, steps: int, warmup: int):
    """Wall ms per call (events on torch's current stream) after warmup."""
    import numpy as np
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(st)
        fn()
        b.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ms = np.array([a.elapsed_time(b) for a, b in ev])
    return wall, ms


def _err(got, ref):
    import numpy as np
    return {"rel_rms": float(np.sqrt(np.mean((got - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))),
            "max_abs": float(np.max(np.abs(got - ref)))}


def _line(metric, value, steps, warmup, ms, dtype, workload, **extra):
    d = {"metric": metric, "value": value, "unit": "points*realizations/s", "n_gpus": 1,
         "steps": steps, "warmup": warmup, "ms_per_step": ms, "higher_is_better": True,
         "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
         "config": {"workload": workload}}
    d.update(extra)
    return d


def run_cfg1(args):
    """cfg 1: Shinozuka on 20,000 points in [0,10]^3, Nmc 4, seed 1, REF_COMPAT,
    the reference's own CPU case; FP64 output compared point-wise."""
    import numpy as np
    import torch

    import randomfield_amd as R
    from oracle import oracle as O
    from randomfield_amd import workloads as W
    x = W.cfg1_points(20000, 1)
    ctx = R.Context(0)
    sptr = torch.cuda.current_stream().cuda_stream
    base = dict(nDim=3, method=R.SHINOZUKA, mode=R.REF_COMPAT, Nmc=4, xRange=(10, 10, 10))
    _, phU = R.reference_draws(R.make_params(**base), R.reference_seed(1))
    po = O.make_plan(3, O.SHINOZUKA, O.REF_COMPAT, [10, 10, 10])
    k = O.kpoints(po)
    sk = O.skvec(po, k)
    t = time.perf_counter()
    ref = O.shinozuka(po, x, k, sk, phU, threads=_threads())
    cpu_s = time.perf_counter() - t
    dx = torch.from_numpy(x).cuda()
    work = 20000 * 4
    forms = {}
    for name, form, dt in (("fp64", R.FORM_FP64, torch.float64), ("fp32", R.FORM_FP32, torch.float32),
                           ("mfma_f16", R.FORM_MFMA_F16, torch.float32)):
        plan = R.FieldPlan(ctx, R.make_params(form=form, **base), None, phU)
        out = torch.empty((4, 20000), dtype=dt, device="cuda")
        wall, ms = _timed(lambda: plan.execute(dx, out, stream=sptr), args.steps, args.warmup)
        forms[name] = {"value": work / (wall * 1e-3), "ms_per_step": wall,
                       "kernel_ms": float(ms.mean()), **_err(out.cpu().numpy().astype(np.float64), ref)}
        plan.close()
    ctx.close()
    return _line("points x realizations / s (cfg1 Shinozuka, FP64 form, point-wise vs FP64 oracle)",
                 forms["fp64"]["value"], args.steps, args.warmup, forms["fp64"]["ms_per_step"], "f64",
                 "cfg1: Shinozuka, 20,000 uniform points in [0,10]^3, N_k 10,976, Nmc 4, REF_COMPAT",
                 forms=forms,
                 cpu_baseline={"value": work / cpu_s, "unit": "points*realizations/s", "cores": _threads(),
                               "kind": "port", "seconds": cpu_s,
                               "sample": "the whole cfg1 workload (FP64 oracle, reference order)"})


def run_cfg2(args):
    """cfg 2: isotropic, 128^3 grid spacing 0.1, corrL 0.5, lognormal (avg 1,
    var 0.25), Nmc 64, seed 7; FP32 direct (per-realization directions), BF16
    and FP16 hi/lo MFMA (shared directions), and the grid form (per-realization
    directions, FP16 hi/lo MFMA); FP32 output with the fused map."""
    import numpy as np
    import torch

    import randomfield_amd as R
    from oracle import oracle as O
    from randomfield_amd import workloads as W
    x = W.grid_points((128, 128, 128), (0.1, 0.1, 0.1))
    N, Nmc = x.shape[0], 64
    ctx = R.Context(0)
    sptr = torch.cuda.current_stream().cuda_stream
    base = dict(nDim=3, method=R.ISOTROPIC, mode=R.FIXED, Nmc=Nmc, corrL=(0.5, 0.5, 0.5),
                xRange=(12.7, 12.7, 12.7), margiFirst=R.FOM_LOGNORMAL, fieldAvg=1.0, fieldVar=0.25)
    _, U = R.reference_draws(R.make_params(**base), R.reference_seed(7))
    dx = torch.from_numpy(x).cuda()
    work = N * Nmc
    sel = np.random.default_rng(0).choice(N, 32, replace=False)
    po = O.make_plan(3, O.ISOTROPIC, O.FIXED, [25.4] * 3)
    k = O.kpoints(po)
    sk = O.skvec(po, k)
    kN = po.kNTotal
    Us = U.copy()
    Us[:, : 2 * kN] = U[0, : 2 * kN]
    forms = {}
    for name, form, shared, uref in (("fp32_direct", R.FORM_FP32, 0, U),
                                     ("mfma_bf16_shared", R.FORM_MFMA_BF16, 1, Us),
                                     ("mfma_f16_shared", R.FORM_MFMA_F16, 1, Us)):
        plan = R.FieldPlan(ctx, R.make_params(form=form, isoShared=shared, **base), None, U)
        out = torch.empty((Nmc, N), dtype=torch.float32, device="cuda")
        wall, ms = _timed(lambda: plan.execute(dx, out, stream=sptr), args.steps, args.warmup)
        ref = O.marginal(2, 1.0, 0.25, O.isotropic(po, x[sel] / 0.5, k, sk, uref))
        got = out[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
        forms[name] = {"value": work / (wall * 1e-3), "ms_per_step": wall, "kernel_ms": float(ms.mean()),
                       **_err(got, ref)}
        plan.close()
    # per-realization directions on the grid form: each term factorises over
    # the axes, one real-part split-FP16 MFMA GEMM per realization over the
    # shells, operands formed from per-axis phase tables
    plan = R.FieldPlan(ctx, R.make_params(form=R.FORM_FP32, isoShared=0, **base), None, U)
    out = torch.empty((Nmc, N), dtype=torch.float32, device="cuda")
    wall, ms = _timed(lambda: plan.execute_grid((0.0, 0.0, 0.0), (0.1, 0.1, 0.1), (128, 128, 128), out,
                                                stream=sptr), args.steps, args.warmup)
    ref = O.marginal(2, 1.0, 0.25, O.isotropic(po, x[sel] / 0.5, k, sk, U))
    got = out[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
    forms["grid_f16_per_realization"] = {"value": work / (wall * 1e-3), "ms_per_step": wall,
                                         "kernel_ms": float(ms.mean()), **_err(got, ref)}
    plan.close()
    # the marginal map alone: 8 B per FP32 value (read + write), HBM-bound
    out = torch.randn((Nmc, N), dtype=torch.float32, device="cuda") * 0.1
    wall, ms = _timed(lambda: R.check(R.lib().rf_marginal_device(ctx.handle, R.FOM_LOGNORMAL, 1.0, 0.25,
                                                                 R._ptr(out), out.numel(), R.DTYPE_F32,
                                                                 R.C.c_void_p(sptr))), 5, 1)
    gbs = 8.0 * out.numel() / (ms.mean() * 1e-3) / 1e9
    t = time.perf_counter()
    O.isotropic(po, x[:2000] / 0.5, k, sk, U, threads=_threads())
    cpu_s = time.perf_counter() - t
    ctx.close()
    best = max(forms.values(), key=lambda f: f["value"])
    return _line("points x realizations / s (cfg2 isotropic + lognormal)", best["value"], args.steps,
                 args.warmup, best["ms_per_step"], "f32",
                 "cfg2: isotropic, 128^3 grid spacing 0.1, corrL 0.5, 286 shells, Nmc 64, lognormal",
                 forms=forms,
                 marginal_map={"bound": "hbm", "achieved_GBps": gbs, "peak_GBps": 8000.0,
                               "frac": gbs / 8000.0, "kernel_ms": float(ms.mean())},
                 cpu_baseline={"value": 2000 * Nmc / cpu_s, "unit": "points*realizations/s",
                               "cores": _threads(), "kind": "port", "seconds": cpu_s,
                               "sample": "first 2,000 grid points, all realizations (FP64 oracle)"})


def run_cfg4(args):
    """cfg 4: 512x512x256 grid (spacing 0.25), 4x4x2 uneven subdomains with
    7-point overlaps, Shinozuka per subdomain on its own lattice (seed 100 +
    field index), Nmc 16, FP32.  A step draws fresh realizations for all 32
    subdomains (new phases from each field's device stream folded into the
    coefficients), generates them (three grouped launches) and blends them
    with the variance-preserving weights, replayed as one captured graph whose
    draw counter advances on every replay.  The re-evaluation of frozen
    realizations (no redraw) is reported beside it."""
    import numpy as np
    import torch

    import randomfield_amd as R
    from oracle import oracle as O
    from randomfield_amd import workloads as W
    groupN, parts, nOvl, step, Nmc = (512, 512, 256), (4, 4, 2), 7, (0.25, 0.25, 0.25), 16
    origins, sizes, masks = W.cfg4_layout(groupN, parts, nOvl)
    ctx = R.Context(0)
    stream = torch.cuda.Stream()
    sptr = stream.cuda_stream
    plans, fields, geo = [], [], []
    for s in range(len(sizes)):
        n = [int(v) for v in sizes[s]]
        o = [origins[s][a] * step[a] for a in range(3)]
        ext = [(n[a] - 1) * step[a] for a in range(3)]
        plans.append(R.FieldPlan(ctx, R.make_params(nDim=3, method=R.SHINOZUKA, mode=R.FIXED,
                                                    form=R.FORM_FP32, Nmc=Nmc, xRange=ext,
                                                    seed=100 + s + 1)))
        fields.append(torch.empty((Nmc, int(np.prod(n))), dtype=torch.float32, device="cuda"))
        geo.append((o, n, ext))
    g = torch.empty((Nmc, int(np.prod(groupN))), dtype=torch.float32, device="cuda")
    batch = R.GridBatch(ctx, plans, [x[0] for x in geo], [step] * len(geo), [x[1] for x in geo], fields,
                        R.DTYPE_F32)
    bplan = R.BlendPlan(ctx, groupN, Nmc, origins, sizes, [nOvl] * 3, masks, R.FIXED, fields, R.DTYPE_F32)
    # tile kinds per contraction (the autotuner) before any capture: the
    # graphs keep the launch shapes of the time they were captured
    kinds = batch.autotune(sptr)
    fresh = R.Graph(ctx, sptr)
    with fresh:
        batch.redraw(R.DRAW_NEXT, sptr)
        batch.execute(sptr)
        bplan.execute(g, sptr)
    replay = R.Graph(ctx, sptr)
    with replay:
        batch.execute(sptr)
        bplan.execute(g, sptr)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t0 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, np.array([a.elapsed_time(b) for a, b in ev])

    wall_f, ms_f = timed(lambda: fresh.launch(sptr), args.steps, args.warmup)
    wall_r, ms_r = timed(lambda: replay.launch(sptr), args.steps, args.warmup)
    _, ms_d = timed(lambda: batch.redraw(R.DRAW_NEXT, sptr), args.steps, 1)
    _, ms_g = timed(lambda: batch.execute(sptr), args.steps, 1)
    _, ms_b = timed(lambda: bplan.execute(g, sptr), args.steps, 1)
    nglob = float(np.prod(groupN))
    nsub = float(sum(np.prod(s) for s in sizes))
    # algorithmic FLOPs of the three contractions per step (8 per complex
    # MAC, 4 per real-part MAC; sign-block fold: K = 2 nk1, 2 nk2, nk3)
    flop_gen = 0.0
    ncoef = 0.0
    for pl, (o_, n_, e_) in zip(plans, geo):
        k1, k2, k3 = [int(v) for v in pl.info.kNStep]
        ncoef += Nmc * pl.info.kNTotal
        flop_gen += Nmc * (8.0 * (2 * k1) * (2 * k2 * k3) * n_[0] + 8.0 * (2 * k2) * k3 * n_[0] * n_[1]
                           + 4.0 * k3 * n_[0] * n_[1] * n_[2])
    work = nglob * Nmc
    bytes_blend = (nsub + nglob) * Nmc * 4.0
    gbs = bytes_blend / (ms_b.mean() * 1e-3) / 1e9
    # accuracy: draw 0 (the plans' own realizations), 24 points of subdomain 0
    batch.redraw(0, sptr)
    batch.execute(sptr)
    torch.cuda.synchronize()
    info = plans[0].info
    du = torch.empty(Nmc * info.kNTotal, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 101, 2, Nmc * info.kNTotal, du)
    torch.cuda.synchronize()
    phU = du.cpu().numpy().reshape(Nmc, info.kNTotal)
    o, n, ext = geo[0]
    po = O.make_plan(3, O.SHINOZUKA, O.FIXED, ext)
    k = O.kpoints(po)
    sk = O.skvec(po, k)
    xs = W.grid_points(n, step, o)
    sel = np.random.default_rng(0).choice(xs.shape[0], 24, replace=False)
    t = time.perf_counter()
    ref = O.shinozuka(po, xs[sel], k, sk, phU, threads=_threads())
    cpu_s = time.perf_counter() - t
    got = fields[0][:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
    fresh.close()
    replay.close()
    # the blend fused into the last contraction: points one subdomain covers
    # go straight to the group grid, the blend forms the overlaps only
    batch.fuse_blend(bplan, g)
    fused_fresh = R.Graph(ctx, sptr)
    with fused_fresh:
        batch.redraw(R.DRAW_NEXT, sptr)
        batch.execute(sptr)
        bplan.execute(g, sptr)
    fused_replay = R.Graph(ctx, sptr)
    with fused_replay:
        batch.execute(sptr)
        bplan.execute(g, sptr)
    wall_ff, ms_ff = timed(lambda: fused_fresh.launch(sptr), args.steps, args.warmup)
    wall_fr, ms_fr = timed(lambda: fused_replay.launch(sptr), args.steps, args.warmup)
    _, ms_fg = timed(lambda: batch.execute(sptr), args.steps, 1)
    _, ms_fb = timed(lambda: bplan.execute(g, sptr), args.steps, 1)
    fused_fresh.close()
    fused_replay.close()
    batch.fuse_blend(bplan, None)
    bplan.close()
    batch.close()
    for pl in plans:
        pl.close()
    ctx.close()
    traffic = None
    tf = ROOT / "profiles" / "traffic_cfg4_step.json"
    if tf.exists():
        traffic = json.loads(tf.read_text())
        traffic["source"] = "profiles/traffic_cfg4_step.json (rocprofv3 --pmc of bench.py --config 4)"
    best_fused = wall_ff <= wall_f
    step_wall = wall_ff if best_fused else wall_f
    # every first contraction on 64x160 tiles: those form the fresh
    # coefficients themselves and the redraw only advances the draw state
    formed = int(kinds[0]) == 5
    return _line("points x realizations / s (cfg4 localization: fresh realizations, generation + blend)",
                 work / (step_wall * 1e-3), args.steps, args.warmup, step_wall, "f32",
                 "cfg4: 512x512x256 grid, 4x4x2 uneven subdomains, 7-point overlap, Nmc 16, "
                 f"N_k {info.kNTotal} (subdomain 0); per step: new phases for every field (device stream, "
                 + ("coefficients formed by the first contraction" if formed else "folded into the coefficients")
                 + "), grid form in grouped launches, variance-preserving blend"
                 + (" fused into the last contraction" if best_fused else "") + "; one captured graph per step",
                 step_kernel_ms=float((ms_ff if best_fused else ms_f).mean()),
                 tile_kinds=kinds,
                 separate_blend={"ms_per_step": wall_f, "kernel_ms": float(ms_f.mean()),
                                 "replay_ms_per_step": wall_r, "generation_ms": float(ms_g.mean()),
                                 "blend_ms": float(ms_b.mean())},
                 fused_blend={"ms_per_step": wall_ff, "kernel_ms": float(ms_ff.mean()),
                              "replay_ms_per_step": wall_fr, "generation_ms": float(ms_fg.mean()),
                              "overlap_blend_ms": float(ms_fb.mean())},
                 frozen_replay={"ms_per_step": wall_fr if best_fused else wall_r,
                                "kernel_ms": float((ms_fr if best_fused else ms_r).mean()),
                                "value": work / ((wall_fr if best_fused else wall_r) * 1e-3),
                                "note": "the same realizations re-evaluated (no redraw)"},
                 redraw_ms=float(ms_d.mean()), generation_ms=float(ms_g.mean()), blend_ms=float(ms_b.mean()),
                 redraw={"coefficients_per_step": ncoef, "kernel_ms": float(ms_d.mean()),
                         "note": ("the draw state advanced by one thread; the first contraction forms each "
                                  "coefficient (Philox uniform, FP32 sincos of 2 pi U, amplitude) while staging it"
                                  if formed else
                                  "one grouped launch over all fields: Philox uniforms, FP32 sincos of 2 pi U, "
                                  "fold; 8 B written per coefficient (FP32 complex)")},
                 generation_roofline={"bound": "mfma", "unit": "TFLOP/s", "peak": 2500.0,
                                      "achieved": flop_gen / (ms_g.mean() * 1e-3) / 1e12,
                                      "frac": flop_gen / (ms_g.mean() * 1e-3) / 1e12 / 2500.0,
                                      "executed_frac": 3 * flop_gen / (ms_g.mean() * 1e-3) / 1e12 / 2500.0,
                                      "flop_per_step": flop_gen, "kernel_ms": float(ms_g.mean()),
                                      "note": "FP16 hi/lo split: 3 MFMA products per algorithmic one"},
                 blend_roofline={"bound": "hbm", "achieved_GBps": gbs, "peak_GBps": 8000.0,
                                 "frac": gbs / 8000.0, "kernel_ms": float(ms_b.mean()),
                                 "bytes_per_launch": bytes_blend},
                 output_bytes_per_step=nglob * Nmc * 4.0,
                 traffic=traffic,
                 accuracy={**_err(got, ref), "sample": "24 points of subdomain 0 x 16 realizations (draw 0)"},
                 cpu_baseline={"value": 24 * Nmc / cpu_s, "unit": "points*realizations/s",
                               "cores": _threads(), "kind": "port", "seconds": cpu_s,
                               "sample": f"24 points x 16 realizations x {info.kNTotal} wavevectors"})


def run_cfg5(args):
    """cfg 5: serving stream.  2,000 requests (seed 23) arriving as a Poisson
    process at args.rate per second: n log-uniform in [1e3, 1e6] points
    uniform in a box of side 10 corrL, corrL U[0.2, 3]^3, Nmc U{1..128},
    method 1 or 2 (50/50), marginal gaussian or lognormal (avg 1, var 0.25).
    Points are resident on the device before the clock starts; the server
    admits by memory budget and batches continuously; a consumer releases
    each field once it is ready."""
    import threading

    import numpy as np
    import torch

    import randomfield_amd as R
    from oracle import oracle as O
    from randomfield_amd.serving import Server
    n_req = args.requests
    rng = np.random.default_rng(23)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n_req))
    ns = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n_req)).astype(np.int64)
    corrLs = rng.uniform(0.2, 3.0, (n_req, 3))
    nmcs = rng.integers(1, 129, n_req)
    methods = np.where(rng.random(n_req) < 0.5, R.ISOTROPIC, R.SHINOZUKA)
    margis = np.where(rng.random(n_req) < 0.5, R.FOM_GAUSSIAN, R.FOM_LOGNORMAL)
    seeds = rng.integers(1, 1 << 40, n_req)
    ctx = R.Context(0)
    xs = []
    for i in range(n_req):
        g = torch.Generator(device="cuda")
        g.manual_seed(int(seeds[i]))
        xs.append(torch.rand((int(ns[i]), 3), dtype=torch.float64, device="cuda", generator=g)
                  * torch.tensor(10.0 * corrLs[i], dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    srv = Server(ctx, hbm_budget=args.budget_gib << 30, max_inflight=3, max_batch=64)
    # warm-up: one request of each method outside the trace
    for m in (R.SHINOZUKA, R.ISOTROPIC):
        rid = srv.submit(xs[0][:2000], m, 8, (1.0, 1.0, 1.0), 1)
        srv.wait(rid)
        srv.release(rid)
    # capacity: the whole stream submitted at once (saturated server)
    t = time.perf_counter()
    burst = [srv.submit(xs[i], int(methods[i]), int(nmcs[i]), corrLs[i], int(seeds[i]), int(margis[i]), 1.0, 0.25)
             for i in range(n_req)]
    for rid in burst:
        srv.wait(rid)
        srv.release(rid)
    burst_s = time.perf_counter() - t
    ids = [None] * n_req
    lat = np.zeros(n_req)
    done = np.zeros(n_req, dtype=bool)

    def consumer():
        for i in range(n_req):
            while ids[i] is None:
                time.sleep(20e-6)
            res = srv.wait(ids[i])
            lat[i] = res.tDone - res.tSubmit
            done[i] = res.status == 0
            srv.release(ids[i])

    th = threading.Thread(target=consumer)
    th.start()
    t0 = time.perf_counter()
    for i in range(n_req):
        while time.perf_counter() - t0 < arrivals[i]:
            time.sleep(20e-6)
        ids[i] = srv.submit(xs[i], int(methods[i]), int(nmcs[i]), corrLs[i], int(seeds[i]), int(margis[i]),
                            1.0, 0.25)
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    work = float(np.sum(ns * nmcs))
    # CPU baseline: the FP64 oracle on the first 500 points of the first
    # request of each method (all its realizations)
    cpu_t, cpu_w = 0.0, 0.0
    for m in (R.SHINOZUKA, R.ISOTROPIC):
        i = int(np.nonzero(methods == m)[0][0])
        xh = xs[i][:500].cpu().numpy() / corrLs[i]
        po = O.make_plan(3, m, O.FIXED, [10.0, 10.0, 10.0])
        k = O.kpoints(po)
        U = np.random.default_rng(0).random((int(nmcs[i]), po.kNTotal * (4 if m == R.ISOTROPIC else 1)))
        t = time.perf_counter()
        if m == R.SHINOZUKA:
            O.shinozuka(po, xh, k, O.skvec(po, k), U, threads=_threads())
        else:
            O.isotropic(po, xh, k, O.skvec(po, k), U, threads=_threads())
        cpu_t += time.perf_counter() - t
        cpu_w += 500.0 * int(nmcs[i])
    ctx.close()
    return _line("points x realizations / s (cfg5 serving stream, p50/p99 latency)", work / wall, 1, 0,
                 wall * 1e3, "f16+f32",
                 f"cfg5: {n_req} Poisson requests at {args.rate}/s, n 1e3..1e6, Nmc 1..128, methods 1/2, "
                 "device-resident points, continuous batching",
                 latency_ms={"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3),
                             "max": float(lat.max() * 1e3)},
                 capacity={"value": float(np.sum(ns * nmcs)) / burst_s, "seconds": burst_s,
                           "note": "same requests submitted at once"},
                 requests={"completed": int(done.sum()), "batches": int(st.batches),
                           "max_batch": int(st.maxBatchSeen), "max_inflight": int(st.maxInflightSeen),
                           "arena_peak_GiB": st.arenaPeak / 2**30, "arena_GiB": st.arenaCapacity / 2**30,
                           "offered_s": float(arrivals[-1]), "makespan_s": wall},
                 cpu_baseline={"value": cpu_w / cpu_t, "unit": "points*realizations/s", "cores": _threads(),
                               "kind": "port", "seconds": cpu_t,
                               "sample": "500 points of the first request of each method, all its realizations"})


def run_cfg6(args):
    """Method 4 (gen_Std_Gauss_FFT): 128^3 grid, spacing 0.1 corrL, Nmc 8,
    DCT-III tables through the grid-form MFMA contractions; FP32 output;
    realization 0 checked against the FP64 oracle.  CPU reference point:
    scipy's FFT-based DCT-III of the same coefficients (the reference calls
    FFTW), one realization."""
    import numpy as np
    import scipy.fft
    import torch

    import randomfield_amd as R
    from oracle import oracle as O
    gn, step, Nmc = (128, 128, 128), (0.1, 0.1, 0.1), 8
    xr = [(n - 1) * s for n, s in zip(gn, step)]
    kN = int(np.prod(gn))
    ctx = R.Context(0)
    sptr = torch.cuda.current_stream().cuda_stream
    u = np.random.default_rng(4).random((Nmc, 3 * kN - 1))
    p = R.make_params(nDim=3, method=R.FFT, mode=R.FIXED, form=R.FORM_FP32, Nmc=Nmc, xRange=xr, gridN=gn)
    plan = R.FieldPlan(ctx, p, None, u)
    out = torch.empty((Nmc, kN), dtype=torch.float32, device="cuda")
    wall_r, ms_r = _timed(lambda: plan.execute_grid((0.0, 0.0, 0.0), step, gn, out, stream=sptr), args.steps,
                          args.warmup)
    po = O.fft_plan(3, xr, gn)
    sk = O.skvec(po, O.kpoints(po))
    ref = O.fft(po, sk, u[:1], step, 1)[0]
    got = out[0].cpu().numpy().astype(np.float64)
    # a step with fresh realizations: new uniforms from the device stream, the
    # DCT coefficients and their fold re-formed, then the three contractions
    draw = [0]

    def fresh():
        draw[0] += 1
        plan.redraw(draw[0], stream=sptr)
        plan.execute_grid((0.0, 0.0, 0.0), step, gn, out, stream=sptr)

    wall, ms = _timed(fresh, args.steps, args.warmup)
    coef = (u[0, :kN] - 0.5) * np.sqrt(sk) * np.cos(2.0 * O.PI_REF * u[0, 2 * kN - 1:])
    t = time.perf_counter()
    scipy.fft.dctn(coef.reshape(gn[::-1]), type=3, workers=_threads())
    cpu_s = time.perf_counter() - t
    plan.close()
    ctx.close()
    work = kN * Nmc
    flop = Nmc * 8.0 * (gn[0] ** 2 * gn[1] * gn[2] + gn[0] * gn[1] ** 2 * gn[2] + 0.5 * gn[0] * gn[1] * gn[2] ** 2)
    return _line("points x realizations / s (method 4, FFT/DCT-III grid)", work / (wall * 1e-3), args.steps,
                 args.warmup, wall, "f32", "method 4: 128^3 grid spacing 0.1 corrL, Nmc 8, DCT-III per axis, "
                 "fresh realizations every step",
                 kernel_ms=float(ms.mean()),
                 reevaluation={"ms_per_step": wall_r, "kernel_ms": float(ms_r.mean()),
                               "note": "same realizations: the plan's folded coefficients reused"},
                 grid_form={"bound": "mfma", "peak": 2500.0, "unit": "TFLOP/s",
                            "achieved": flop / (ms_r.mean() * 1e-3) / 1e12,
                            "frac": flop / (ms_r.mean() * 1e-3) / 1e12 / 2500.0,
                            "executed_frac": 3 * flop / (ms_r.mean() * 1e-3) / 1e12 / 2500.0,
                            "note": "complex contractions on FP16 MFMA with hi/lo split operands (3 products), "
                                    "8 FLOP per complex MAC (algorithmic), last stage real part only"},
                 accuracy=_err(got, ref),
                 cpu_reference={"value": kN / cpu_s, "unit": "points*realizations/s", "cores": _threads(),
                                "kind": "scipy.fft.dctn type 3 (FFT-based, as FFTW REDFT01)",
                                "seconds": cpu_s, "sample": "one realization"})


def measured_peaks():
    """Runs lib/rf_peaks (a child process) and returns its JSON, or None."""
    exe = ROOT / "randomfield_amd" / "lib" / "rf_peaks"
    if not exe.exists():
        return None
    import subprocess
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--direct-steps", type=int, default=2)
    ap.add_argument("--cpu-points", type=int, default=10000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--rate", type=float, default=600.0, help="cfg 5: arrivals per second")
    ap.add_argument("--requests", type=int, default=2000, help="cfg 5: requests in the stream")
    ap.add_argument("--budget-gib", type=int, default=64, help="cfg 5: server memory budget")
    ap.add_argument("--peaks", action="store_true", help="only print the measured unit ceilings (JSON)")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the cfg1 / cfg4 / cfg5 records of the default line")
    ap.add_argument("--cpu-points-1core", type=int, default=400,
                    help="points of the single-core CPU baseline sample")
    ap.add_argument("--config", type=int, default=3,
                    help="benchmark configuration 1..6 (3, the headline workload, by default; "
                         "6 is method 4, the FFT method)")
    args = ap.parse_args()
    if args.peaks:
        print(json.dumps(measured_peaks()), flush=True)
        return 0
    if args.config != 3 and not args.calibrate_copy:
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            return 2
        runners = {1: run_cfg1, 2: run_cfg2, 4: run_cfg4, 5: run_cfg5, 6: run_cfg6}
        if args.config not in runners:
            print(f"unknown configuration {args.config}", file=sys.stderr)
            return 2
        print(json.dumps(runners[args.config](args)), flush=True)
        return 0
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4}))
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "documented (points and realizations shard with no exchange) but not run.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import randomfield_amd as R
    from randomfield_amd import workloads as W

    x = W.gaussian_mixture_points(N_POINTS, 11)
    ctx = R.Context(0)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream
    params = dict(nDim=3, method=R.RANDOMIZATION, mode=R.FIXED, kSampling=R.KSAMP_DENSITY,
                  nKDensity=N_K, Nmc=N_MC, corrL=CORR_L, xRange=(1, 1, 1), seed=11)
    p = R.make_params(form=R.FORM_MFMA_F16, **params)
    plan = R.FieldPlan(ctx, p)
    dx = torch.from_numpy(x).cuda()
    out = torch.empty((N_MC, N_POINTS), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    # launch shapes: each timed briefly, then the library's own autotuner
    # picks the one used below
    variants_ms = []
    for v in range(R.MFMA_VARIANTS):
        plan.set_variant(v)
        _, vms = _timed(lambda: plan.execute(dx, out, stream=sptr), 3, 1)
        variants_ms.append(float(vms.mean()))
    chosen = plan.autotune(dx, out, stream=sptr)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        plan.execute(dx, out, stream=sptr)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        plan.execute(dx, out, stream=sptr)
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = np.array([a.elapsed_time(b) for a, b in ev])          # per launch, ms
    ms_step = wall * 1e3 / args.steps
    work = float(N_POINTS) * N_MC
    value = work / (wall / args.steps)
    flop_alg = 4.0 * N_K * work                                  # 2 FLOP x 2 N_k MACs
    achieved = flop_alg / (kms.mean() * 1e-3)

    # plain FP32 direct sum on the same workload (skipped with --direct-steps 0)
    direct = None
    out_d = None
    dplan = None
    if args.direct_steps > 0:
        pd = R.make_params(form=R.FORM_FP32, **params)
        dplan = R.FieldPlan(ctx, pd)
        out_d = torch.empty_like(out)
        dplan.execute(dx, out_d, stream=sptr)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.direct_steps):
            dplan.execute(dx, out_d, stream=sptr)
        torch.cuda.synchronize()
        d_ms = (time.perf_counter() - t1) * 1e3 / args.direct_steps
        direct_value = work / (d_ms * 1e-3)
        direct = {"value": direct_value, "ms_per_step": d_ms, "mfma_speedup": value / direct_value}

    # accuracy of sampled points against the FP64 oracle (same device uniforms)
    from oracle import oracle as O
    info = plan.info
    du = torch.empty(info.nKUniform + N_MC * N_K, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, 11, 1, info.nKUniform, du)
    R.device_uniforms(ctx, 11, 2, N_MC * N_K, du[info.nKUniform:])
    torch.cuda.synchronize()
    hu = du.cpu().numpy()
    kU, phU = hu[: info.nKUniform], hu[info.nKUniform:].reshape(N_MC, N_K)
    po = O.make_plan(3, O.RANDOMIZATION, O.FIXED, [1, 1, 1], ksamp=O.KDENSITY, nKDensity=N_K)
    kt = O.kpoints(po, kU)
    skt = O.skvec(po, kt)
    sel = np.random.default_rng(0).choice(N_POINTS, 32, replace=False)
    ref = O.shinozuka(po, x[sel] / np.array(CORR_L), kt, skt, phU)
    got = out[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
    rms = float(np.sqrt(np.mean(ref ** 2)))
    acc = {"rel_rms": float(np.sqrt(np.mean((got - ref) ** 2)) / rms),
           "max_abs": float(np.max(np.abs(got - ref))),
           "sample": "32 random points x 256 realizations vs FP64 oracle"}
    if out_d is not None:
        got_d = out_d[:, torch.from_numpy(sel).cuda()].cpu().numpy().astype(np.float64)
        acc["direct_fp32_rel_rms"] = float(np.sqrt(np.mean((got_d - ref) ** 2)) / rms)

    # PCIe-inclusive rate, reported beside the HBM-resident value and never as
    # it: host points in, host field out through pinned buffers, one launch
    try:
        hx = torch.empty((N_POINTS, 3), dtype=torch.float64, pin_memory=True)
        hx.copy_(torch.from_numpy(x))
        hout = torch.empty((N_MC, N_POINTS), dtype=torch.float32, pin_memory=True)
        torch.cuda.synchronize()
        t4 = time.perf_counter()
        dx.copy_(hx, non_blocking=True)
        plan.execute(dx, out, stream=sptr)
        hout.copy_(out, non_blocking=True)
        torch.cuda.synchronize()
        pcie_s = time.perf_counter() - t4
        pcie = {"value": work / pcie_s, "ms": pcie_s * 1e3, "unit": "points*realizations/s",
                "note": "48 MB of points host->device, one launch, the 2.05 GB FP32 field device->host "
                        "(pinned host buffers)"}
        del hx, hout
    except (RuntimeError, MemoryError) as e:        # no pinned memory: the headline still prints
        pcie = {"error": f"{type(e).__name__}: {e}"}

    # CPU baseline: the oracle (the reference algorithm, FP64, libm cos) on a
    # bounded point sample of the same workload, all host threads
    cpu = None
    if not args.no_cpu:
        threads = os.cpu_count() or 1
        try:
            threads = len(os.sched_getaffinity(0))
        except AttributeError:
            pass
        threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
        xs = x[: args.cpu_points] / np.array(CORR_L)
        t2 = time.perf_counter()
        O.shinozuka(po, xs, kt, skt, phU, threads=threads)
        cs = time.perf_counter() - t2
        cpu = {"value": args.cpu_points * N_MC / cs, "unit": "points*realizations/s", "cores": threads,
               "kind": "port", "seconds": cs,
               "sample": f"first {args.cpu_points} of the 2,000,000 points, all 256 realizations and "
                         f"4096 wavevectors (FP64 oracle, reference operation order)"}
        # one core: the reference runs serially per MPI rank
        n1 = args.cpu_points_1core
        t3 = time.perf_counter()
        O.shinozuka(po, xs[:n1], kt, skt, phU, threads=1)
        c1 = time.perf_counter() - t3
        cpu["single_core"] = {"value": n1 * N_MC / c1, "cores": 1, "seconds": c1,
                              "sample": f"first {n1} points, all 256 realizations and 4096 wavevectors"}

    traffic = None
    traffic_src = None
    tf = ROOT / "profiles" / "traffic_cfg3_mfma_f16.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        traffic_src = ("profiles/traffic_cfg3_mfma_f16.json: an earlier rocprofv3 --pmc run of this "
                       "kernel (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch), not measured in this run")

    # the f16 MFMA ceiling measured on this device with operands that change
    # every instruction (the datasheet peak assumes a clock the chip does not
    # hold under that load); executed work = 3 x algorithmic (hi/lo split)
    ceiling = None
    pk = measured_peaks()
    if pk is not None:
        ceiling = {"mfma_f16_toggling_TFLOPs": pk["mfma_f16_toggling_TFLOPs"],
                   "mfma_f16_constant_TFLOPs": pk["mfma_f16_constant_TFLOPs"],
                   "executed_frac": 3 * achieved / 1e12 / pk["mfma_f16_toggling_TFLOPs"]}
    line = {
        "metric": "points x realizations / s (cfg3 randomization, <=1e-5 rel RMS vs FP64)",
        "value": value,
        "unit": "points*realizations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": "cfg3: randomization, 2,000,000 Gaussian-mixture points, N_k 4096 "
                               "(spectral-density samples), Nmc 256, corrL (2,2,0.5)",
                   "form": "MFMA FP16 hi/lo, FP32 accumulate, FP64 k.x reduction",
                   "points": N_POINTS, "realizations": N_MC, "wavevectors": N_K},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_F16_DENSE / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_F16_DENSE, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel_ms": float(kms.mean()),
                     "flop_per_launch": flop_alg, "executed_mfma_flop_per_launch": 3 * flop_alg,
                     "variant": chosen, "variants_kernel_ms": variants_ms,
                     "measured_ceiling": ceiling},
        "latency_ms": {"p50": float(np.percentile(kms, 50)), "p99": float(np.percentile(kms, 99))},
        "direct_fp32": direct,
        "accuracy": acc,
        "pcie_inclusive": pcie,
        "cpu_baseline": cpu,
    }
    plan.close()
    if dplan is not None:
        dplan.close()
    ctx.close()
    del dx, out, out_d
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_extras:
        # the other workloads, each on its own context, in the same run
        extras = {}
        for name, fn in (("cfg1", run_cfg1), ("cfg4", run_cfg4), ("cfg5", run_cfg5)):
            t = time.perf_counter()
            try:
                extras[name] = fn(args)
            except Exception as e:                       # recorded, the headline still prints
                extras[name] = {"error": f"{type(e).__name__}: {e}"}
            extras[name]["bench_seconds"] = time.perf_counter() - t
            torch.cuda.synchronize()
            torch.cuda.empty_cache()
            print(f"[bench] {name} done in {extras[name]['bench_seconds']:.1f} s", file=sys.stderr, flush=True)
        line["other_configs"] = extras
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
