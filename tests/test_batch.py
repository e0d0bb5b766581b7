"""Batched persistent launch: several models' volumes in one launch per
method give exactly the per-model results (same kernels bodies, same voxel
groups, same Philox draws), and the batch C-ABI validates its items."""
import numpy as np
import pytest

from tests.helpers import dev, host

# (model, voxels, voxel-group size): NODDI_Bingham_2Fans at its tuned 64
# lanes (the column-wise Jacobian, J^T J in LDS)
BATCH = [("DiffusionTensor", 300, 4), ("NODDI_Bingham_2Fans", 160, 64), ("Ball_3_Sticks", 250, 4)]


def test_default_batch_unit_is_compiled():
    from cudimot_amd import capi
    units = capi.batch_units()
    assert ["DiffusionTensor", "NODDI_Bingham_2Fans", "Ball_3_Sticks"] in units.values()


def test_batch_argument_errors_without_gpu():
    from cudimot_amd import capi
    a, b = capi.Model("Ball_1_Stick"), capi.Model("DiffusionTensor")
    fake = dict(meas=_Fake((10, 64)), params=8, cfp=8)
    with pytest.raises(capi.CdmError, match="no batch unit"):
        capi.batch_levmar([dict(model=a, **fake)])
    with pytest.raises(capi.CdmError, match="share a model handle"):
        capi.batch_levmar([dict(model=b, **fake), dict(model=b, **fake)])
    with pytest.raises(capi.CdmError, match="at least one"):
        capi.batch_levmar([])
    # the Rician likelihood needs every item's tau samples (checked before any device work)
    items = [dict(model=capi.Model(name), meas=_Fake((10, 64)), params=8, cfp=8, fixp=8, samples=8)
             for name, _, _ in BATCH]
    with pytest.raises(capi.CdmError, match="needs tau_samples"):
        capi.batch_mcmc(items, nburn=10, njumps=10, sampleevery=5, rician=True)
    # the per-item kernel threshold is process-wide; negative values are refused
    with pytest.raises(capi.CdmError, match="blocks per CU"):
        capi.batch_set_fill(-1)
    capi.batch_set_fill(0)
    capi.batch_set_fill(8)


class _Fake:
    """Stands in for a device tensor in host-side argument checks."""

    def __init__(self, shape):
        self.shape = shape

    def data_ptr(self):
        return 8


def _volumes(seed):
    import torch
    from cudimot_amd import capi, synth
    out = []
    for k, (name, n, L) in enumerate(BATCH):
        cfp = synth.model_acquisition(name)
        t = synth.truth_for(name, n, seed + k)
        m = capi.Model(name)
        fixp = torch.full((n, 1), 1000.0, device="cuda") if m.fixp_tsize else None
        clean = torch.empty((n, cfp.shape[0]), device="cuda")
        m.predicted_signal(dev(t.reshape(n, -1, 1)), dev(cfp), fixp, clean)
        c = host(clean)
        rng = np.random.default_rng(seed + 10 + k)
        s0 = t[:, 0] if name != "NODDI_Bingham_2Fans" else np.full(n, 1000.0)
        sig = (s0 / 20.0)[:, None]
        meas = np.sqrt((c + sig * rng.normal(size=c.shape)) ** 2 + (sig * rng.normal(size=c.shape)) ** 2)
        start = t.copy()
        start[:, 1:] *= 1.0 + 0.05 * rng.normal(size=start[:, 1:].shape)  # a perturbed start
        if name == "NODDI_Bingham_2Fans":
            start = synth.clip_to_bounds(name, start)
        out.append(dict(name=name, L=L, meas=dev(meas), cfp=dev(cfp), fixp=fixp, start=start, n=n))
    return out


def _models(vols, method):
    from cudimot_amd import capi
    ms = []
    for v in vols:
        m = capi.Model(v["name"])
        m.set_group(method, v["L"])
        ms.append(m)
    return ms


@pytest.mark.gpu
def test_batch_levmar_equals_per_model_levmar():
    import torch
    from cudimot_amd import capi
    vols = _volumes(400)
    ref_models, bat_models = _models(vols, capi.LEVMAR), _models(vols, capi.LEVMAR)
    ref, items = [], []
    for v, m, mb in zip(vols, ref_models, bat_models):
        p = dev(v["start"])
        c = torch.empty(v["n"], device="cuda")
        it = torch.empty(v["n"], dtype=torch.int32, device="cuda")
        m.levmar(v["meas"], p, v["cfp"], v["fixp"], max_iters=100, cost=c, iters=it)
        ref.append((p, c, it))
        pb = dev(v["start"])
        cb = torch.empty(v["n"], device="cuda")
        ib = torch.empty(v["n"], dtype=torch.int32, device="cuda")
        items.append(dict(model=mb, meas=v["meas"], params=pb, cfp=v["cfp"], fixp=v["fixp"], cost=cb, iters=ib))
    capi.batch_levmar(items, max_iters=100)
    capi.sync()
    for (p, c, it), b in zip(ref, items):
        np.testing.assert_array_equal(host(p), host(b["params"]))
        np.testing.assert_array_equal(host(c), host(b["cost"]))
        np.testing.assert_array_equal(host(it), host(b["iters"]))


@pytest.mark.gpu
@pytest.mark.parametrize("blocks,graph", [(0, False), (20, True)])
def test_batch_mcmc_equals_per_model_mcmc(blocks, graph):
    import torch
    from cudimot_amd import capi
    vols = _volumes(500)
    ref_models, bat_models = _models(vols, capi.MCMC), _models(vols, capi.MCMC)
    ref, items = [], []
    off = 0
    for v, m, mb in zip(vols, ref_models, bat_models):
        P = m.nparams
        p = dev(v["start"])
        s = torch.empty((v["n"], P, 4), device="cuda")
        m.mcmc(v["meas"], p, v["cfp"], v["fixp"], s, None, nburn=60, njumps=40, sampleevery=10, seed=77,
               voxel_offset=off)
        ref.append((p, s))
        pb = dev(v["start"])
        sb = torch.empty((v["n"], P, 4), device="cuda")
        items.append(dict(model=mb, meas=v["meas"], params=pb, cfp=v["cfp"], fixp=v["fixp"], samples=sb,
                          voxel_offset=off))
        off += v["n"]
    capi.batch_mcmc(items, nburn=60, njumps=40, sampleevery=10, seed=77, jumps_per_launch=blocks, use_graph=graph)
    capi.sync()
    for (p, s), b in zip(ref, items):
        np.testing.assert_array_equal(host(s), host(b["samples"]))
        np.testing.assert_array_equal(host(p), host(b["params"]))


@pytest.mark.gpu
@pytest.mark.parametrize("fill", [8, 0])
def test_batch_item_that_fills_the_gpu_runs_its_own_kernels(fill):
    """A DiffusionTensor item of 140,000 voxels (2,188 blocks of its 4-lane
    voxel groups, over 8 per CU) goes through the per-model kernels at its
    tuned launch shape, the two small items through the persistent launch;
    LM and MCMC results equal the per-model calls' either way, and the MCMC
    jump-block graphs land in the big item's handle and in the batch
    workspace respectively.  With cdm_batch_set_fill(0) every item goes
    through the persistent launch, with the same results."""
    import torch
    from cudimot_amd import capi, synth
    vols = _volumes(700)
    n = 140000
    cfp = synth.model_acquisition("DiffusionTensor")
    t = synth.truth_for("DiffusionTensor", n, 71)
    m0 = capi.Model("DiffusionTensor")
    clean = torch.empty((n, cfp.shape[0]), device="cuda")
    m0.predicted_signal(dev(t.reshape(n, -1, 1)), dev(cfp), None, clean)
    rng = np.random.default_rng(72)
    c = host(clean)
    sig = (t[:, 0] / 20.0)[:, None]
    meas = np.sqrt((c + sig * rng.normal(size=c.shape)) ** 2 + (sig * rng.normal(size=c.shape)) ** 2)
    start = t.copy()
    start[:, 1:] *= 1.0 + 0.05 * rng.normal(size=start[:, 1:].shape)
    vols[0] = dict(name="DiffusionTensor", L=None, meas=dev(meas), cfp=dev(cfp), fixp=None, start=start, n=n)
    ref_models = [capi.Model(v["name"]) for v in vols]
    bat_models = [capi.Model(v["name"]) for v in vols]
    for v, a, b in zip(vols, ref_models, bat_models):
        for h in (a, b):
            if v["L"]:  # the small items at fixed voxel groups (the persistent launch's are chosen for its total)
                h.set_group(capi.LEVMAR, v["L"])
                h.set_group(capi.MCMC, v["L"])
    assert ref_models[0].get_group(capi.MCMC, cfp.shape[0], n) == 4
    ref, items = [], []
    off = 0
    for v, m, mb in zip(vols, ref_models, bat_models):
        P = m.nparams
        p = dev(v["start"])
        cost = torch.empty(v["n"], device="cuda")
        m.levmar(v["meas"], p, v["cfp"], v["fixp"], max_iters=50, cost=cost)
        s = torch.empty((v["n"], P, 2), device="cuda")
        m.set_mcmc_blocks(20, use_graph=True)
        m.mcmc(v["meas"], p, v["cfp"], v["fixp"], s, None, nburn=40, njumps=20, sampleevery=10, seed=5,
               voxel_offset=off)
        ref.append((p, cost, s))
        items.append(dict(model=mb, meas=v["meas"], params=dev(v["start"]), cfp=v["cfp"], fixp=v["fixp"],
                          cost=torch.empty(v["n"], device="cuda"), samples=torch.empty((v["n"], P, 2), device="cuda"),
                          voxel_offset=off))
        off += v["n"]
    before = capi.batch_graph_stats()
    capi.batch_set_fill(fill)
    try:
        capi.batch_levmar(items, max_iters=50)
        capi.batch_mcmc(items, nburn=40, njumps=20, sampleevery=10, seed=5, jumps_per_launch=20, use_graph=True)
        capi.sync()
    finally:
        capi.batch_set_fill(8)
    for (p, cost, s), b in zip(ref, items):
        np.testing.assert_array_equal(host(cost), host(b["cost"]))
        np.testing.assert_array_equal(host(s), host(b["samples"]))
        np.testing.assert_array_equal(host(p), host(b["params"]))
    assert bat_models[1].graph_stats() == (0, 0) and bat_models[2].graph_stats() == (0, 0)
    if fill:
        assert bat_models[0].graph_stats()[0] == 1  # the burn-in's two blocks, recorded on the handle
    else:
        assert bat_models[0].graph_stats() == (0, 0)
    assert sum(capi.batch_graph_stats()) == sum(before) + 1


@pytest.mark.gpu
@pytest.mark.parametrize("blocks,graph", [(0, False), (25, True)])
def test_batch_rician_mcmc_equals_per_model_mcmc(blocks, graph):
    """The Rician likelihood (MCMC.cu:799-847) in the batched launch: samples,
    tau samples and final parameters equal the per-model Rician MCMC's."""
    import torch
    from cudimot_amd import capi
    vols = _volumes(600)
    ref_models, bat_models = _models(vols, capi.MCMC), _models(vols, capi.MCMC)
    ref, items = [], []
    off = 0
    for v, m, mb in zip(vols, ref_models, bat_models):
        P = m.nparams
        p = dev(v["start"])
        s = torch.empty((v["n"], P, 5), device="cuda")
        tau = torch.empty((v["n"], 5), device="cuda")
        m.mcmc(v["meas"], p, v["cfp"], v["fixp"], s, tau, nburn=50, njumps=50, sampleevery=10, rician=True, seed=91,
               voxel_offset=off)
        ref.append((p, s, tau))
        pb = dev(v["start"])
        sb = torch.empty((v["n"], P, 5), device="cuda")
        tb = torch.empty((v["n"], 5), device="cuda")
        items.append(dict(model=mb, meas=v["meas"], params=pb, cfp=v["cfp"], fixp=v["fixp"], samples=sb,
                          tau_samples=tb, voxel_offset=off))
        off += v["n"]
    capi.batch_mcmc(items, nburn=50, njumps=50, sampleevery=10, seed=91, jumps_per_launch=blocks, use_graph=graph,
                    rician=True)
    capi.sync()
    for (p, s, tau), b in zip(ref, items):
        assert np.all(np.isfinite(host(tau))) and np.all(host(tau) > 0)
        np.testing.assert_array_equal(host(s), host(b["samples"]))
        np.testing.assert_array_equal(host(tau), host(b["tau_samples"]))
        np.testing.assert_array_equal(host(p), host(b["params"]))


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_mcmc_jump_blocks_do_not_change_draws(graph):
    """A phase split into jump blocks (optionally replayed from a HIP graph)
    gives the draws of the single launch."""
    import torch
    from cudimot_amd import capi
    v = _volumes(600)[2]
    out = []
    for blocks in (0, 15):
        m = capi.Model(v["name"])
        m.set_group(capi.MCMC, v["L"])
        if blocks:
            m.set_mcmc_blocks(blocks, use_graph=graph)
        p = dev(v["start"])
        s = torch.empty((v["n"], m.nparams, 5), device="cuda")
        m.mcmc(v["meas"], p, v["cfp"], v["fixp"], s, None, nburn=50, njumps=50, sampleevery=10, seed=5)
        capi.sync()
        out.append((host(p), host(s)))
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][0], out[1][0])


@pytest.mark.gpu
def test_mcmc_graph_cached_and_replayed():
    """A blocked MCMC call with use_graph records each phase's launches once;
    a second call with the same arguments replays the kept graphs (no new
    capture) and gives bit-identical draws, equal to the unblocked launch;
    other arguments (a new seed) record new graphs.  Prints the host time
    of issuing one launch per jump against replaying the recorded graphs."""
    import time
    import torch
    from cudimot_amd import capi
    v = _volumes(700)[0]
    m = capi.Model(v["name"])
    m.set_group(capi.MCMC, v["L"])
    nb, nj, se = 30, 40, 10

    # the same output buffers in every call: a graph bakes its arguments in
    p = dev(v["start"])
    s = torch.empty((v["n"], m.nparams, nj // se), device="cuda")
    start = dev(v["start"])

    def run(model, seed=5):
        p.copy_(start)
        model.mcmc(v["meas"], p, v["cfp"], v["fixp"], s, None, nburn=nb, njumps=nj, sampleevery=se, seed=seed)
        capi.sync()
        return host(p), host(s)
    ref = run(m)
    m.set_mcmc_blocks(1, use_graph=True)  # one launch per jump
    first = run(m)
    assert m.graph_stats() == (2, 0)
    second = run(m)
    assert m.graph_stats() == (2, 2)
    for got in (first, second):
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
    run(m, seed=6)
    assert m.graph_stats() == (4, 2)
    # host cost of submitting 70 launches, against replaying two graphs
    mu = capi.Model(v["name"])
    mu.set_group(capi.MCMC, v["L"])
    mu.set_mcmc_blocks(1, use_graph=False)
    t_issue, t_replay = [], []
    for _ in range(5):
        capi.sync()
        t0 = time.perf_counter()
        run(mu)
        t_issue.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        run(m, seed=6)
        t_replay.append(time.perf_counter() - t0)
    print("MCMC, 70 jump launches: issued %.3f ms, replayed from the cached graphs %.3f ms (median of 5)" %
          (1e3 * np.median(t_issue), 1e3 * np.median(t_replay)))


@pytest.mark.gpu
def test_batch_mcmc_graph_replayed_and_async():
    """Two identical batched MCMC calls with graphs: the second replays the
    workspace's graphs, both are bit-identical; the calls return without
    synchronising the stream (results read after capi.sync)."""
    import torch
    from cudimot_amd import capi
    vols = _volumes(800)
    models = _models(vols, capi.MCMC)
    outs = []
    c0, r0 = capi.batch_graph_stats()
    for rep in range(2):
        items = []
        off = 0
        for v, m in zip(vols, models):
            items.append(dict(model=m, meas=v["meas"], params=dev(v["start"]), cfp=v["cfp"], fixp=v["fixp"],
                              samples=torch.empty((v["n"], m.nparams, 4), device="cuda"), voxel_offset=off))
            off += v["n"]
        capi.batch_mcmc(items, nburn=40, njumps=40, sampleevery=10, seed=9, jumps_per_launch=10, use_graph=True)
        outs.append(items)
    capi.sync()
    c1, r1 = capi.batch_graph_stats()
    assert (c1 - c0, r1 - r0) == (2, 2)
    for a, b in zip(*outs):
        np.testing.assert_array_equal(host(a["samples"]), host(b["samples"]))
        np.testing.assert_array_equal(host(a["params"]), host(b["params"]))


def test_batch_lm_wide_item_falls_back_like_per_model_launch():
    """A NODDI_Bingham_2Fans item with 400 measurements at 64 LM lanes needs
    more LDS than a CU has (its J rows); the batch preparation falls back to
    16 lanes as the per-model launch does instead of refusing the item.
    Checked on the host: preparation passes, and only the device work that
    follows can fail here."""
    from cudimot_amd import capi
    m = capi.Model("NODDI_Bingham_2Fans")
    m.set_group(capi.LEVMAR, 64)
    dt, d3 = capi.Model("DiffusionTensor"), capi.Model("Ball_3_Sticks")
    items = [dict(model=m, meas=_Fake((1000, 400)), params=8, cfp=8, fixp=8),
             dict(model=dt, meas=_Fake((10, 64)), params=8, cfp=8),
             dict(model=d3, meas=_Fake((10, 280)), params=8, cfp=8)]
    try:
        capi.batch_levmar(items)
    except capi.CdmError as e:
        assert "does not fit" not in str(e), e
