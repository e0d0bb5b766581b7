"""Full-size runs of the measured configurations, checked through properties
that do not depend on the volume size, with oracle spot checks on random
subsets (the oracle cannot fit 500k voxels in test time):

  * NODDI_Watson, 500k voxels, M=99: the reference pipeline's GridSearch +
    LM recipe (Pipeline_NODDI_Watson.sh:185-246);
  * Ball_2_Sticks, 500k voxels, M=280: GridSearch -> LM -> MCMC (1000 burn-in,
    1250 jumps, a sample every 25);
  * DiffusionTensor / NODDI_Bingham_2Fans / Ball_3_Sticks, 400k voxels each:
    one batched launch per method, LM (200 iterations) -> MCMC (500 burn-in,
    1250 jumps), LM checked against the oracle on a random subset.

Properties: every output finite and inside the bounds; LM never ends above
its start; GridSearch picks the oracle's combination on every sampled voxel;
LM from the same start reaches the oracle's cost (Watson); MCMC draws of a
voxel range re-run alone (streams addressed by the voxel's global index) and
per-model launches on a slice of a batched volume are bit-identical to the
full runs."""
import numpy as np
import pytest

from oracle import bind
from tests.helpers import dev, host

pytestmark = pytest.mark.gpu


def _volume(model, n, seed, snr, s0=1000.0):
    """Noisy signals on the device for synthetic truths (host [n][P])."""
    import torch
    from cudimot_amd import capi, synth
    cfp = synth.model_acquisition(model)
    t = synth.truth_for(model, n, seed)
    m = capi.Model(model)
    fx = torch.full((n, 1), s0, device="cuda") if m.fixp_tsize else None
    clean = torch.empty((n, cfp.shape[0]), device="cuda")
    m.predicted_signal(dev(t.reshape(n, -1, 1)), dev(cfp), fx, clean)
    scale = dev(t[:, :1]) if not m.fixp_tsize else torch.full((n, 1), s0, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    sig = scale / snr
    n1 = torch.randn(clean.shape, device="cuda", generator=g)
    n2 = torch.randn(clean.shape, device="cuda", generator=g)
    meas = torch.sqrt((clean + sig * n1) ** 2 + (sig * n2) ** 2).contiguous()
    return cfp, t, meas, fx


def _in_bounds(spec, p, tol=1e-5):
    p = np.asarray(p)
    for i, bt in enumerate(spec["bound_types"]):
        lo, hi = np.float32(spec["bmin"][i]), np.float32(spec["bmax"][i])
        span = max(1.0, abs(float(lo)), abs(float(hi)))
        if bt in (1, 3):
            assert np.all(p[..., i] >= lo - tol * span), (i, p[..., i].min(), lo)
        if bt in (2, 3):
            assert np.all(p[..., i] <= hi + tol * span), (i, p[..., i].max(), hi)


def _sse(m, meas, p, cfp, fx):
    import torch
    n = meas.shape[0]
    s = torch.empty_like(meas)
    m.predicted_signal(p.reshape(n, -1, 1).contiguous(), dev(cfp), fx, s)
    return ((s.double() - meas.double()) ** 2).sum(1)


def test_watson_500k_gridsearch_lm():
    import torch
    from cudimot_amd import capi, pipelines
    model, n = "NODDI_Watson", 500_000
    cfp, t, meas, _ = _volume(model, n, seed=41, snr=25.0)
    b0 = torch.as_tensor(pipelines.b0_mask(cfp), device="cuda")
    s0 = meas[:, b0].mean(1, keepdim=True).contiguous()  # FixP S0 = mean b=0 (Pipeline_NODDI_Watson.sh:156-171)
    frame = pipelines.tensor_frame(meas, cfp)
    kept = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    p, cost = pipelines.watson_gridsearch_lm(meas, cfp, s0, frame=frame, kept=kept)
    capi.sync()
    ph, ch, kh = host(p), host(cost), kept.cpu().numpy()
    assert np.all(np.isfinite(ph)) and np.all(np.isfinite(ch))
    assert np.all(kh >= 0)
    _in_bounds(capi.Model(model).spec(), ph)
    # the recipe again on a random subset, step by step with the oracle
    rng = np.random.default_rng(3)
    idx = np.sort(rng.choice(n, 128, replace=False))
    o = bind.OracleModel(model)
    gs = capi.Model(model)
    gs.set_fixed("3,4")
    gp, grid = gs.load_grid(pipelines.WATSON_GRID, is_text=True)
    sp = {k: np.asarray(v) for k, v in gs.spec().items() if k != "init_provided"}
    meas_s = host(meas)[idx].astype(np.float32).astype(np.float64)
    s0_s = host(s0)[idx]
    cfp32 = cfp.astype(np.float32).astype(np.float64)
    p0 = np.tile(np.asarray(sp["p_init"], dtype=np.float32), (len(idx), 1)).astype(np.float64)
    p0[:, 3:5] = host(frame)[idx, 0:2]
    p_gs, kept_ref = o.gridsearch(sp, meas_s, p0, cfp32, s0_s, np.asarray(gp), np.asarray(grid))
    np.testing.assert_array_equal(kh[idx], kept_ref)
    fr = capi.Model(model)
    fr.set_fixed("2,3,4")
    spf = {k: np.asarray(v) for k, v in fr.spec().items() if k != "init_provided"}
    spa = {k: np.asarray(v) for k, v in capi.Model(model).spec().items() if k != "init_provided"}
    p1, _, _ = o.levmar(spf, meas_s, p_gs.astype(np.float32).astype(np.float64), cfp32, s0_s, max_iters=200)
    p2, _, c_ref = o.levmar(spa, meas_s, p1.astype(np.float32).astype(np.float64), cfp32, s0_s, max_iters=200)
    c_gpu = ((o.signal(ph[idx].astype(np.float64), cfp32, s0_s) - meas_s) ** 2).sum(1)
    agree = np.mean(np.abs(c_gpu - c_ref) / c_ref <= 1e-3)
    print("Watson 500k: oracle cost agreement on the subset", agree)
    assert agree >= 0.95, agree


def test_ball2sticks_500k_gridsearch_lm_mcmc():
    import torch
    from cudimot_amd import capi
    from tests.test_gpu_parity import GRIDS
    model, n = "Ball_2_Sticks", 500_000
    cfp, t, meas, _ = _volume(model, n, seed=42, snr=20.0)
    b0 = torch.as_tensor(cfp[:, 3] == 0, device="cuda")
    m = capi.Model(model)
    spec = m.spec()
    p = dev(np.tile(np.asarray(spec["p_init"], dtype=np.float32), (n, 1)))
    p[:, 0] = meas[:, b0].mean(1)
    gp, grid = m.load_grid(GRIDS[model], is_text=True)
    kept = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    m.gridsearch(meas, p, dev(cfp), None, kept=kept)
    p_gs = p.clone()
    cost = torch.empty(n, device="cuda")
    m.levmar(meas, p, dev(cfp), None, max_iters=200, cost=cost)
    p_lm = p.clone()
    ns = 1250 // 25
    samples = torch.empty((n, 8, ns), device="cuda")
    m.mcmc(meas, p, dev(cfp), None, samples, None, nburn=1000, njumps=1250, sampleevery=25, seed=8219)
    capi.sync()
    sh, ch = host(samples), host(cost)
    assert np.all(np.isfinite(sh)) and np.all(np.isfinite(ch))
    _in_bounds(spec, np.moveaxis(sh, 2, 1))
    # LM never ends above its start (its own FP64 cost vs the FP32-signal SSE at the start)
    c_gs = _sse(m, meas, p_gs, cfp, None).cpu().numpy()
    assert np.all(ch <= c_gs * (1 + 1e-5) + 1e-3), np.max(ch / c_gs)
    # GridSearch: the oracle's combination on a random subset
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(n, 96, replace=False))
    o = bind.OracleModel(model)
    sp = {k: np.asarray(v) for k, v in spec.items() if k != "init_provided"}
    p0 = np.tile(np.asarray(sp["p_init"], dtype=np.float32), (len(idx), 1)).astype(np.float64)
    p0[:, 0] = host(p_gs)[idx, 0]
    meas_s = host(meas)[idx].astype(np.float32).astype(np.float64)
    _, kept_ref = o.gridsearch(sp, meas_s, p0, cfp.astype(np.float32).astype(np.float64), None, np.asarray(gp),
                               np.asarray(grid))
    np.testing.assert_array_equal(kept.cpu().numpy()[idx], kept_ref)
    # MCMC streams follow the voxel's global index: a voxel range run alone
    # (same group size) reproduces the full run's draws bit for bit
    a, k = 123_456, 4096
    m2 = capi.Model(model)
    m2.set_group(capi.MCMC, m.get_group(capi.MCMC, cfp.shape[0], n))
    ps = p_lm[a:a + k].clone()
    ss = torch.empty((k, 8, ns), device="cuda")
    m2.mcmc(meas[a:a + k].contiguous(), ps, dev(cfp), None, ss, None, nburn=1000, njumps=1250, sampleevery=25,
            seed=8219, voxel_offset=a)
    capi.sync()
    np.testing.assert_array_equal(host(ss), sh[a:a + k])


def test_batch_400k_three_models_lm_mcmc():
    import torch
    from cudimot_amd import capi, synth
    models = [("DiffusionTensor", 51), ("NODDI_Bingham_2Fans", 52), ("Ball_3_Sticks", 53)]
    n = 400_000
    items, vols, handles = [], [], []
    offset = 0
    for name, seed in models:
        cfp, t, meas, fx = _volume(name, n, seed=seed, snr=20.0)
        rng = np.random.default_rng(seed)
        start = t.copy()
        start[:, 1:] *= 1.0 + 0.05 * rng.normal(size=start[:, 1:].shape)
        start = synth.clip_to_bounds(name, start)
        h = capi.Model(name)
        L_lm, L_mc = h.get_group(capi.LEVMAR, cfp.shape[0], n), h.get_group(capi.MCMC, cfp.shape[0], n)
        h.set_group(capi.LEVMAR, L_lm)
        h.set_group(capi.MCMC, L_mc)
        P = h.nparams
        it = dict(model=h, meas=meas, params=dev(start), cfp=dev(cfp), fixp=fx,
                  cost=torch.empty(n, device="cuda"), iters=torch.empty(n, dtype=torch.int32, device="cuda"),
                  samples=torch.empty((n, P, 50), device="cuda"), voxel_offset=offset)
        items.append(it)
        vols.append((name, cfp, start, meas, fx, L_lm, L_mc, offset))
        offset += n
    capi.batch_levmar(items, max_iters=200)
    capi.sync()
    lm_out = [host(it["params"]) for it in items]
    # LM against the oracle from the same starts on a random subset of each
    # model: the GPU not in worse minima more often than better (+5%), and
    # most voxels at the oracle's cost
    rng = np.random.default_rng(5)
    for (name, cfp, start, meas, fx, L_lm, L_mc, off), p_lm in zip(vols, lm_out):
        idx = np.sort(rng.choice(n, 64, replace=False))
        o = bind.OracleModel(name)
        sp = {k: np.asarray(v) for k, v in capi.Model(name).spec().items() if k != "init_provided"}
        meas_s = host(meas)[idx].astype(np.float32).astype(np.float64)
        cfp32 = cfp.astype(np.float32).astype(np.float64)
        fx_s = host(fx)[idx] if fx is not None else None
        p_ref, _, _ = o.levmar(sp, meas_s, start[idx].astype(np.float32).astype(np.float64), cfp32, fx_s,
                               max_iters=200)
        cost = lambda q: ((o.signal(q.astype(np.float32).astype(np.float64), cfp32, fx_s) - meas_s) ** 2).sum(1)
        lr = np.log(cost(p_lm[idx]) / cost(p_ref))  # both after FixConstraintsLM, float-rounded
        worse, better, same = np.mean(lr > 1e-3), np.mean(lr < -1e-3), np.mean(np.abs(lr) <= 1e-3)
        print(name, "400k batched LM vs oracle on 64 voxels: worse", worse, "better", better, "same", same)
        # NODDI_Bingham_2Fans: the oracle against itself after arithmetic-level
        # perturbations agrees on ~72% of the voxels (test_levmar_2fans_unbiased)
        assert worse <= better + (0.1 if name == "NODDI_Bingham_2Fans" else 0.05) + 1e-12, (name, worse, better)
        assert same >= (0.6 if name == "NODDI_Bingham_2Fans" else 0.75), (name, same)
    capi.batch_mcmc(items, nburn=500, njumps=1250, sampleevery=25, seed=8219)
    capi.sync()
    a, k = 200_000, 4096
    for it, (name, cfp, start, meas, fx, L_lm, L_mc, off), p_lm in zip(items, vols, lm_out):
        s = host(it["samples"])
        assert np.all(np.isfinite(s)) and np.all(np.isfinite(p_lm)), name
        _in_bounds(it["model"].spec(), np.moveaxis(s, 2, 1))
        # per-model launches on a slice give the batched rows bit for bit
        h = capi.Model(name)
        h.set_group(capi.LEVMAR, L_lm)
        h.set_group(capi.MCMC, L_mc)
        ps = dev(start[a:a + k])
        fxs = fx[a:a + k].contiguous() if fx is not None else None
        h.levmar(meas[a:a + k].contiguous(), ps, dev(cfp), fxs, max_iters=200)
        capi.sync()
        np.testing.assert_array_equal(host(ps), p_lm[a:a + k], err_msg=name)
        ss = torch.empty((k, h.nparams, 50), device="cuda")
        h.mcmc(meas[a:a + k].contiguous(), ps, dev(cfp), fxs, ss, None, nburn=500, njumps=1250, sampleevery=25,
               seed=8219, voxel_offset=off + a)
        capi.sync()
        np.testing.assert_array_equal(host(ss), s[a:a + k], err_msg=name)
