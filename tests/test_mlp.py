"""Learned initialiser: the numpy oracle pinned by finite differences and
closed forms (CPU), and the BF16-MFMA network kernels against the oracle
(GPU): forward, loss gradient, decoding, training."""
import numpy as np
import pytest

from oracle import mlp_ref as R
from tests.helpers import dev, host

BINGHAM_HEAD = [
    dict(kind=R.SIGMOID, param=0, param2=-1, lo=0.0, hi=1.0, weight=1.0, vmin=0.0, vmax=1.0),
    dict(kind=R.SIGMOID, param=1, param2=-1, lo=0.0, hi=1.0, weight=1.0, vmin=0.0, vmax=1.0),
    dict(kind=R.SIGMOID_LOG, param=2, param2=-1, lo=0.1, hi=64.0, weight=1.0, vmin=0.11, vmax=63.9),
    dict(kind=R.SIGMOID_RATIO, param=3, param2=2, lo=0.0, hi=1.0, weight=1.0, vmin=0.011, vmax=31.9),
    dict(kind=R.AXIS, param=4, param2=5, lo=0.0, hi=1.0, weight=2.0, vmin=0.0, vmax=0.0),
]


def bingham_truth(n, seed):
    from cudimot_amd import synth
    return synth.truth_for("NODDI_Bingham", n, seed)


def random_layers(shapes, rng, scale=1.0):
    return [(rng.normal(size=s) * scale / np.sqrt(s[1]), rng.normal(size=s[0]) * 0.1) for s in shapes]


# ------------------------------------------------------------------ CPU: oracle pins

def test_oracle_gradient_matches_finite_differences():
    """The FP64 backward pass equals central differences of the FP64 loss."""
    rng = np.random.default_rng(0)
    shapes = [(16, 11), (16, 16), (7, 16)]
    layers = random_layers(shapes, rng)
    x = rng.uniform(0.1, 1.0, size=(24, 11))
    p = bingham_truth(24, 5)
    loss, g = R.loss_grad(layers, BINGHAM_HEAD, x, p)
    w0 = R.flatten(layers)
    idx = rng.choice(w0.size, 40, replace=False)
    for k in idx:
        h = 1e-6
        wp, wm = w0.copy(), w0.copy()
        wp[k] += h
        wm[k] -= h
        lp, _ = R.loss_grad(R.unflatten(wp, shapes), BINGHAM_HEAD, x, p)
        lm, _ = R.loss_grad(R.unflatten(wm, shapes), BINGHAM_HEAD, x, p)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_oracle_head_decode_inverts_encode():
    """decode(logit(encode(p))) = p for sigmoid entries; the axis decodes to
    the same fibre (up to sign)."""
    p = bingham_truth(200, 9)
    p[:, 3] = np.minimum(p[:, 3], 31.0)
    t = R.encode(BINGHAM_HEAD, p)
    out = t.copy()
    s = np.clip(t[:, :4], 1e-9, 1 - 1e-9)
    out[:, :4] = np.log(s / (1 - s))
    q = R.decode(BINGHAM_HEAD, out, np.zeros_like(p))
    np.testing.assert_allclose(q[:, :4], np.clip(p[:, :4], [0, 0, 0.11, 0.011], [1, 1, 63.9, 31.9]), rtol=1e-6,
                               atol=1e-9)
    n1 = np.stack([np.sin(q[:, 4]) * np.cos(q[:, 5]), np.sin(q[:, 4]) * np.sin(q[:, 5]), np.cos(q[:, 4])], 1)
    n0 = t[:, 4:7]
    np.testing.assert_allclose(np.abs(np.sum(n0 * n1, 1)), 1.0, atol=1e-9)


def test_oracle_loss_zero_at_target():
    """Loss is 0 (up to the clip) at the encoded targets and positive off them."""
    p = bingham_truth(50, 3)
    t = R.encode(BINGHAM_HEAD, p)
    out = t.copy()
    s = np.clip(t[:, :4], 1e-7, 1 - 1e-7)
    out[:, :4] = np.log(s / (1 - s))
    l0, g0 = R.loss_and_dout(BINGHAM_HEAD, out, t)
    assert l0 < 1e-5 and np.abs(g0).max() < 1e-6  # the axis term carries the 1e-6 guard
    l1, _ = R.loss_and_dout(BINGHAM_HEAD, out + 0.3, t)
    assert l1 > 1e-3


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3])
    r = R.round_bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == 1.0
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_mlp_argument_errors_without_gpu():
    from cudimot_amd import capi
    with pytest.raises(capi.CdmError, match="width"):
        capi.Mlp(99, [capi.head_entry(capi.HEAD_SIGMOID, 0)], width=100)
    with pytest.raises(capi.CdmError, match="head entry"):
        capi.Mlp(99, [capi.head_entry(capi.HEAD_SIGMOID_LOG, 0, lo=0.0, hi=1.0)])


# ------------------------------------------------------------------ GPU: kernels vs oracle

def make_net(seed=7, width=256, depth=3):
    from cudimot_amd import capi
    head = [capi.head_entry(e["kind"], e["param"], e["param2"], e["lo"], e["hi"], e["weight"], e["vmin"], e["vmax"])
            for e in BINGHAM_HEAD]
    return capi.Mlp(99, head, width=width, depth=depth, seed=seed)


def bingham_signals(n, seed, snr=None):
    """Normalised NODDI-Bingham signals (S0 = 1) from the GPU forward model."""
    import torch
    from cudimot_amd import capi, synth
    m = capi.Model("NODDI_Bingham")
    cfp = synth.model_acquisition("NODDI_Bingham")
    t = bingham_truth(n, seed)
    out = torch.empty((n, cfp.shape[0]), device="cuda")
    m.predicted_signal(dev(t.reshape(n, 7, 1)), dev(cfp), torch.ones((n, 1), device="cuda"), out)
    capi.sync()
    s = host(out)
    if snr:
        rng = np.random.default_rng(seed + 100)
        s = np.sqrt((s + rng.normal(size=s.shape) / snr) ** 2 + (rng.normal(size=s.shape) / snr) ** 2)
    return t, s.astype(np.float32), cfp


@pytest.mark.gpu
def test_mlp_forward_matches_oracle():
    import torch
    net = make_net()
    t, x, _ = bingham_signals(512, 11, snr=30)
    out = torch.empty((512, net.n_out), device="cuda")
    net.forward(dev(x), None, out)
    layers = R.unflatten(net.get_weights(), net.layer_shapes())
    ref, _ = R.forward(layers, x, bf16=True)
    got = host(out)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-2, err


@pytest.mark.gpu
def test_mlp_loss_gradient_matches_oracle():
    net = make_net(seed=3)
    t, x, _ = bingham_signals(256, 12, snr=20)
    loss, g = net.loss_grad(dev(x), None, dev(t))
    layers = R.unflatten(net.get_weights(), net.layer_shapes())
    lref, gref = R.loss_grad(layers, BINGHAM_HEAD, x, t, bf16=True)
    assert abs(loss - lref) <= 1e-3 * abs(lref), (loss, lref)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    assert rel < 2e-2, rel
    # per layer as well: every weight matrix and bias vector
    k = 0
    for (o, i) in net.layer_shapes():
        for size in (o * i, o):
            a, b = g[k:k + size], gref[k:k + size]
            assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b) + 1e-9, (o, i, size)
            k += size


@pytest.mark.gpu
def test_mlp_init_params_decodes_outputs():
    import torch
    net = make_net(seed=5)
    t, x, _ = bingham_signals(300, 13, snr=25)
    out = torch.empty((300, net.n_out), device="cuda")
    net.forward(dev(x), None, out)
    p = dev(np.full((300, 7), 1.5))
    net.init_params(dev(x), None, p)
    ref = R.decode(BINGHAM_HEAD, host(out).astype(np.float64), np.full((300, 7), 1.5))
    np.testing.assert_allclose(host(p), ref, rtol=2e-5, atol=2e-5)


@pytest.mark.gpu
def test_mlp_training_lowers_loss_and_beats_default_start():
    """Trained on GPU-generated signals, the network's starting points fit
    held-out voxels better than the plug-in's P_init start (the method
    without it), before and after LM."""
    import torch
    from cudimot_amd import capi
    net = make_net(seed=21)
    t, clean, cfp = bingham_signals(65536, 31)
    losses = net.train(dev(clean), None, dev(t), snr=(10, 50), epochs=12, batch=1024, lr=2e-3, seed=4)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.6 * losses[0], losses

    th, xh, _ = bingham_signals(2048, 77, snr=25)
    m = capi.Model("NODDI_Bingham")
    spec = m.spec()
    p_def = np.tile(np.asarray(spec["p_init"], dtype=np.float64), (2048, 1))
    p_net = dev(p_def)
    net.init_params(dev(xh), None, p_net)
    pn = host(p_net)
    # orientation error of the learned start (degrees)
    def axis(a):
        return np.stack([np.sin(a[:, 4]) * np.cos(a[:, 5]), np.sin(a[:, 4]) * np.sin(a[:, 5]), np.cos(a[:, 4])], 1)
    ang = np.degrees(np.arccos(np.clip(np.abs(np.sum(axis(pn) * axis(th), 1)), 0, 1)))
    assert np.median(ang) < 20.0, np.median(ang)
    assert np.median(np.abs(pn[:, 1] - th[:, 1])) < 0.1

    fx = torch.ones((2048, 1), device="cuda")
    cost = {}
    for name, start in (("net", pn), ("default", p_def)):
        p = dev(start)
        c = torch.empty(2048, device="cuda")
        m.levmar(dev(xh), p, dev(cfp), fx, max_iters=200, cost=c)
        capi.sync()
        cost[name] = host(c)
    assert np.median(cost["net"]) < np.median(cost["default"]), (np.median(cost["net"]), np.median(cost["default"]))


@pytest.mark.gpu
def test_mlp_training_graph_equals_step_by_step_launches():
    """cdm_mlp_train replays one captured step graph (its changing arguments
    from a device table) or launches every step's kernels: the same weights
    after three epochs, and training twice in a row (the graph is captured
    anew, the Adam step count carried on) also agrees."""
    from cudimot_amd import capi
    t, clean, _ = bingham_signals(4096, 41)
    weights = {}
    for use_graph in (True, False):
        net = make_net(seed=22)
        net.set_train_graph(use_graph)
        for rep in range(2):
            net.train(dev(clean), None, dev(t), snr=(10, 50), epochs=3, batch=512, lr=2e-3, seed=5 + rep)
            capi.sync()
            weights[(use_graph, rep)] = net.get_weights()
    for rep in range(2):
        np.testing.assert_array_equal(weights[(True, rep)], weights[(False, rep)])
    assert not np.array_equal(weights[(True, 0)], weights[(True, 1)])
