"""--debug=<voxel> (cudimotoptions.h:160-161): GridSearch, LM and MCMC write
records of one voxel's state to a device buffer, which the host reads.  The
records must agree with the methods' own outputs for that voxel: the kept
combination, the start and final LM parameters and the accepted-iteration
count, and the MCMC parameters at every recorded sample."""
import numpy as np
import pytest

from tests import helpers as H


def test_debug_arguments_without_gpu():
    from cudimot_amd import capi
    m = capi.Model("Ball_1_Stick")
    with pytest.raises(capi.CdmError, match="at least one record"):
        m.set_debug(3, 0)
    m.set_debug(-1)  # off: nothing to free, nothing recorded
    assert m.debug_trace(capi.LEVMAR).shape == (0, 9)
    with pytest.raises(capi.CdmError, match="GridSearch, Levenberg-Marquardt and MCMC"):
        m.debug_trace(capi.PREDICTED)


@pytest.mark.gpu
@pytest.mark.parametrize("lanes", [1, 16])
def test_debug_records_follow_the_voxel(lanes):
    import torch
    from cudimot_amd import capi
    from oracle import bind
    model, n, dv = "Ball_1_Stick", 300, 123
    o = bind.OracleModel(model)
    cfp, t, fx, noisy = H.volume(o, model, n, 31)
    m = capi.Model(model)
    for meth in (capi.GRIDSEARCH, capi.LEVMAR, capi.MCMC):
        m.set_group(meth, lanes)
    m.load_grid("search[1]=(0.001,0.002,0.003)\nsearch[2]=(0.2,0.5,0.8)\n", is_text=True)
    m.set_debug(dv, 2048)
    meas, d_cfp, p = H.dev(noisy), H.dev(cfp), H.dev(np.tile(np.asarray(m.spec()["p_init"]), (n, 1)))
    kept = torch.empty(n, dtype=torch.int32, device="cuda")
    m.gridsearch(meas, p, d_cfp, None, kept=kept)
    capi.sync()
    g = m.debug_trace(capi.GRIDSEARCH)
    assert g.shape == (1, 9)
    assert int(g[0, 0]) == int(H.host(kept)[dv])
    np.testing.assert_array_equal(g[0, 4:].astype(np.float32), H.host(p)[dv])
    start = H.host(p)[dv].copy()

    it = torch.empty(n, dtype=torch.int32, device="cuda")
    cost = torch.empty(n, device="cuda")
    m.levmar(meas, p, d_cfp, None, max_iters=30, iters=it, cost=cost)
    capi.sync()
    r = m.debug_trace(capi.LEVMAR)
    print("LM records", r.shape[0], "accepted", int(r[1:, 3].sum()), flush=True)
    np.testing.assert_array_equal(r[0, 4:].astype(np.float32), start)
    np.testing.assert_array_equal(r[:, 0], np.arange(r.shape[0]))
    assert int(r[1:, 3].sum()) == int(H.host(it)[dv])
    np.testing.assert_array_equal(r[-1, 4:].astype(np.float32), H.host(p)[dv])
    acc = r[r[:, 3] > 0]
    assert np.float32(acc[-1, 1]) == H.host(cost)[dv]
    assert np.all(np.diff(acc[:, 1]) < 0)  # accepted attempts lower the cost

    nb, nj, se = 20, 30, 10
    s = torch.empty((n, 5, nj // se), device="cuda")
    m.mcmc(meas, p, d_cfp, None, s, None, nburn=nb, njumps=nj, sampleevery=se, seed=5)
    capi.sync()
    q = m.debug_trace(capi.MCMC)
    assert q.shape == (nb + nj, 9)
    np.testing.assert_array_equal(q[:, 0], np.arange(nb + nj))
    for k in range(nj // se):
        np.testing.assert_array_equal(q[nb + k * se, 4:].astype(np.float32), H.host(s)[dv, :, k])
    np.testing.assert_array_equal(q[-1, 4:].astype(np.float32), H.host(p)[dv])

    # a voxel outside the call's volume records nothing
    m.set_debug(n + 5, 16)
    m.levmar(meas, p, d_cfp, None, max_iters=3)
    capi.sync()
    assert m.debug_trace(capi.LEVMAR).shape[0] == 0
