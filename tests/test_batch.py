"""Batched persistent launch: several models' volumes in one launch per
method give exactly the per-model results (same kernels bodies, same voxel
groups, same Philox draws), and the batch C-ABI validates its items."""
import numpy as np
import pytest

from tests.helpers import dev, host

BATCH = [("DiffusionTensor", 300, 4), ("NODDI_Bingham_2Fans", 160, 16), ("Ball_3_Sticks", 250, 4)]


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
