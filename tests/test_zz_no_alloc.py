"""cdm_mcmc_noalloc: MCMC with its chain state in the library's static device
scratch instead of a device allocation.  Same kernels and draws as
cdm_mcmc, so samples, tau samples and final parameters are bit-identical to
cdm_mcmc's on the same inputs: checked on two calls in a row, the second on
a smaller volume (the scratch keeps the first call's contents, which the
second must not depend on), for a model with per-measurement cache rows
(NODDI_Bingham) at its tuned voxel group, where the rows stay in LDS, and at
one lane per voxel, where a block's 256 voxels need 2 x 99 rows each (about
200 KB) and the rows go to the scratch after the chain states; for a plain
model (Ball_1_Stick), with jump blocks, and with the Rician likelihood.
Sizes beyond the scratch are refused on the host before any device work."""
import numpy as np
import pytest

from tests import helpers as H


def _volume(model, n, seed):
    from cudimot_amd import synth
    from oracle import bind
    o = bind.OracleModel(model)
    cfp, t, fx, noisy = H.volume(o, model, n, seed)
    return cfp, synth.clip_to_bounds(model, t), fx, noisy


def _run(torch, m, entry, cfp, start, fx, noisy, rician, nb=40, nj=60, se=5):
    n = noisy.shape[0]
    p = H.dev(start)
    samples = torch.full((n, m.nparams, nj // se), -1.0, dtype=torch.float32, device="cuda")
    tau = torch.full((n, nj // se), -1.0, dtype=torch.float32, device="cuda") if rician else None
    getattr(m, entry)(H.dev(noisy), p, H.dev(cfp), H.dev(fx), samples, tau, nburn=nb, njumps=nj, sampleevery=se,
                      rician=rician, seed=321, voxel_offset=17)
    return p, samples, tau


@pytest.mark.gpu
@pytest.mark.parametrize("model,rician,lanes,blocks", [("NODDI_Bingham", False, None, 0),
                                                       ("NODDI_Bingham", False, 1, 0),
                                                       ("Ball_1_Stick", False, None, 0),
                                                       ("Ball_1_Stick", False, None, 20),
                                                       ("Ball_1_Stick", True, None, 0)])
def test_mcmc_noalloc_matches_mcmc(model, rician, lanes, blocks):
    import torch
    from cudimot_amd import capi
    assert torch.cuda.is_available()
    print("noalloc", model, "rician", rician, "lanes", lanes, "blocks", blocks, flush=True)
    for n, seed in ((3000, 5), (700, 6)):  # two calls in a row, the second smaller
        cfp, t, fx, noisy = _volume(model, n, seed)
        ref = capi.Model(model)
        m = capi.Model(model)
        for h in (ref, m):
            if lanes:
                h.set_group(capi.MCMC, lanes)
            if blocks:
                h.set_mcmc_blocks(blocks, use_graph=True)
        a = _run(torch, ref, "mcmc", cfp, t, fx, noisy, rician)
        b = _run(torch, m, "mcmc_noalloc", cfp, t, fx, noisy, rician)
        capi.sync()
        for x, y in zip(a, b):
            if x is None:
                continue
            xa, ya = x.cpu().numpy(), y.cpu().numpy()
            assert np.all(np.isfinite(xa))
            np.testing.assert_array_equal(ya, xa)


def test_mcmc_noalloc_refuses_oversized_volume():
    """Checked on the host before any device work (the pointers are never
    dereferenced): 10^6 NODDI_Bingham voxels at 99 measurements need at
    least 256 MB of chain state."""
    from cudimot_amd import capi
    m = capi.Model("NODDI_Bingham")
    L = capi.lib()
    fake = 1 << 20
    r = L.cdm_mcmc_noalloc(m.h, 1000000, 99, fake, fake, fake, fake, 10, 10, 5, 1, 0, 1, 0, fake, None, None)
    assert r != 0
    msg = L.cdm_last_error().decode()
    assert "allocation-free scratch" in msg and "voxel_offset" in msg, msg
