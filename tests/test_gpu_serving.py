"""GPU tests of the serving loop: a served request equals the same field
built by a plan (bit for bit) and the FP64 oracle (within the form's
tolerance); results are the same alone and in a mixed stream; admission
respects the memory budget."""
import numpy as np
import pytest

import randomfield_amd as R
from oracle import oracle as O
from randomfield_amd.serving import Server

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def box_points(n, corrL, seed, side=10.0):
    rng = np.random.default_rng(seed)
    return rng.random((n, 3)) * (side * np.asarray(corrL))


def device_phase_uniforms(ctx, seed, count):
    import torch
    d = torch.empty(count, dtype=torch.float64, device="cuda")
    R.device_uniforms(ctx, seed, 2, count, d)
    ctx.sync()
    return d.cpu().numpy()


@pytest.mark.parametrize("method,Nmc,margi", [(R.SHINOZUKA, 5, R.FOM_NONE), (R.SHINOZUKA, 40, R.FOM_LOGNORMAL),
                                              (R.ISOTROPIC, 3, R.FOM_NONE), (R.ISOTROPIC, 20, R.FOM_GAUSSIAN)])
def test_served_equals_plan_and_oracle(ctx, method, Nmc, margi):
    import torch
    corrL = (0.5, 2.0, 1.0)             # powers of two: box / corrL is exactly 10
    x = box_points(3000, corrL, 1 + Nmc)
    seed = 77 + Nmc
    avg, var = 1.0, 0.25
    with Server(ctx, hbm_budget=1 << 30) as srv:
        rid = srv.submit(torch.from_numpy(x).cuda(), method, Nmc, corrL, seed, margi, avg, var)
        got = srv.read(rid)
        srv.release(rid)
    form = R.FORM_MFMA_F16 if method == R.SHINOZUKA else R.FORM_FP32
    p = R.make_params(nDim=3, method=method, mode=R.FIXED, form=form, Nmc=Nmc, corrL=corrL,
                      xRange=[10 * c for c in corrL], margiFirst=margi, fieldAvg=avg, fieldVar=var, seed=seed)
    plan = R.FieldPlan(ctx, p)
    out = torch.empty((Nmc, x.shape[0]), dtype=torch.float32, device="cuda")
    plan.execute(torch.from_numpy(x).cuda(), out)
    ctx.sync()
    info = plan.info
    plan.close()
    assert np.array_equal(got, out.cpu().numpy())
    # FP64 oracle on the same lattice and device phase draws
    po = O.make_plan(3, method, O.FIXED, [10.0, 10.0, 10.0])
    k = O.kpoints(po)
    U = device_phase_uniforms(ctx, seed, Nmc * info.nPhaseUniform).reshape(Nmc, info.nPhaseUniform)
    xs = x / np.asarray(corrL)
    if method == R.SHINOZUKA:
        ref = O.shinozuka(po, xs, k, O.skvec(po, k), U)
    else:
        ref = O.isotropic(po, xs, k, O.skvec(po, k), U)
    if margi != R.FOM_NONE:
        ref = O.marginal(margi, avg, var, ref)
    e = rel_rms(got.astype(np.float64), ref)
    print(f"[parity] served method {method} Nmc {Nmc}: rel_rms={e:.3e}", flush=True)
    assert e <= 1e-5


def _mixed_requests(rng, n_req):
    reqs = []
    for q in range(n_req):
        corrL = rng.uniform(0.2, 3.0, 3)
        n = int(np.exp(rng.uniform(np.log(500), np.log(40000))))
        reqs.append(dict(x=box_points(n, corrL, 1000 + q), method=int(rng.choice([R.SHINOZUKA, R.ISOTROPIC])),
                         Nmc=int(rng.integers(1, 80)), corrL=corrL, seed=int(rng.integers(1, 1 << 40)),
                         margi=int(rng.choice([R.FOM_GAUSSIAN, R.FOM_LOGNORMAL])), avg=1.0, var=0.25))
    return reqs


def test_alone_vs_mix_identical(ctx):
    """Each result is the same served alone and in a mixed stream (several
    batches in flight, grouped launches of different requests)."""
    import torch
    rng = np.random.default_rng(23)
    reqs = _mixed_requests(rng, 14)
    dev = [torch.from_numpy(r["x"]).cuda() for r in reqs]
    alone = []
    with Server(ctx, hbm_budget=2 << 30, max_inflight=1, max_batch=1) as srv:
        for r, x in zip(reqs, dev):
            rid = srv.submit(x, r["method"], r["Nmc"], r["corrL"], r["seed"], r["margi"], r["avg"], r["var"])
            alone.append(srv.read(rid))
            srv.release(rid)
    with Server(ctx, hbm_budget=2 << 30, max_inflight=3, max_batch=5) as srv:
        ids = [srv.submit(x if i % 3 else r["x"], r["method"], r["Nmc"], r["corrL"], r["seed"], r["margi"],
                          r["avg"], r["var"]) for i, (r, x) in enumerate(zip(reqs, dev))]
        mixed = [srv.read(i) for i in ids]
        batches = {srv.wait(i).batch for i in ids}
        for i in ids:
            srv.release(i)
        st = srv.stats()
    assert len(batches) >= 3 and st.maxBatchSeen > 1
    for q, (a, b) in enumerate(zip(alone, mixed)):
        assert a.shape == b.shape and np.array_equal(a, b), q


def test_admission_budget(ctx):
    """Requests whose total exceeds the budget all complete (queued until
    memory frees up); a request that can never fit is refused."""
    import torch
    budget = 96 << 20
    x = torch.from_numpy(box_points(60000, (1.0, 1.0, 1.0), 5)).cuda()
    with Server(ctx, hbm_budget=budget, max_inflight=2, max_batch=4) as srv:
        ids = [srv.submit(x, R.SHINOZUKA, 100, (1.0, 1.0, 1.0), 10 + i) for i in range(6)]   # ~24 MB out each
        ref = None
        for i, rid in enumerate(ids):
            res = srv.wait(rid)
            assert res.status == 0
            f = srv.read(rid)
            srv.release(rid)
            if i == 0:
                ref = f
            assert np.isfinite(f).all() and abs(float(f.std()) - 1.0) < 0.2
        st = srv.stats()
        assert st.completed == 6 and st.arenaPeak <= st.arenaCapacity
        big = torch.zeros((2_000_000, 3), dtype=torch.float64, device="cuda")
        with pytest.raises(R.RFError) as e:
            srv.submit(big, R.SHINOZUKA, 64, (1.0, 1.0, 1.0), 1)
        assert e.value.code == R.RF_E_NOMEM
    assert ref is not None
