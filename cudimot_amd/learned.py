"""Learned initialiser: output heads per model, training-set generation with
the engine's own forward model, training, and weight files.

The training signals come from the plug-in's Predicted_Signal on the GPU
(cdm_predicted_signal with one sample per voxel) at S0 = 1, so network
inputs are signals divided by S0; Rician noise is drawn afresh every epoch
inside the training call.  No counterpart in the reference.
"""
import os

import numpy as np

from . import capi, synth

WIDTH, DEPTH = 256, 3
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights")


def head(model):
    """Output head of the shipped NODDI plug-ins: bounded scalars through
    sigmoids (kappa on a log scale, beta as a fraction of kappa) and the fibre
    axis through a spherical map (three outputs -> theta, phi).  Bingham's psi
    is left to a short scan (cudimot_amd.pipelines)."""
    H = capi.head_entry
    if model == "NODDI_Bingham":
        return [H(capi.HEAD_SIGMOID, 0, lo=0.0, hi=1.0, vmin=0.0, vmax=1.0),
                H(capi.HEAD_SIGMOID, 1, lo=0.0, hi=1.0, vmin=0.0, vmax=1.0),
                H(capi.HEAD_SIGMOID_LOG, 2, lo=0.1, hi=64.0, vmin=0.11, vmax=63.9),
                H(capi.HEAD_SIGMOID_RATIO, 3, 2, lo=0.0, hi=1.0, vmin=0.011, vmax=31.9),
                H(capi.HEAD_AXIS, 4, 5, weight=2.0)]
    if model == "NODDI_Watson":
        return [H(capi.HEAD_SIGMOID, 0, lo=0.0, hi=1.0, vmin=0.0, vmax=1.0),
                H(capi.HEAD_SIGMOID, 1, lo=0.0, hi=1.0, vmin=0.0, vmax=1.0),
                H(capi.HEAD_SIGMOID_LOG, 2, lo=0.01, hi=64.0, vmin=0.011, vmax=63.9),
                H(capi.HEAD_AXIS, 3, 4, weight=2.0)]
    raise ValueError("no learned-initialiser head for model " + model)


def training_set(model, n, seed, stream=None):
    """Ground truths [n][P] (host) and their noise-free signals at S0 = 1
    (device [n][M]) for the model's acquisition."""
    import torch
    cfp = synth.model_acquisition(model)
    t = synth.truth_for(model, n, seed)
    m = capi.Model(model)
    d_cfp = torch.as_tensor(cfp, dtype=torch.float32).cuda()
    d_t = torch.as_tensor(t, dtype=torch.float32).cuda()
    clean = torch.empty((n, cfp.shape[0]), device="cuda")
    ones = torch.ones((n, 1), device="cuda")
    chunk = 1 << 20
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m.predicted_signal(d_t[a:b].reshape(b - a, -1, 1), d_cfp, ones[a:b], clean[a:b], stream=stream)
    return t, d_t, clean, cfp


def train(model="NODDI_Bingham", n=2_000_000, epochs=20, batch=2048, lr=2e-3, snr=(10.0, 50.0), seed=1234,
          stream=None):
    """Train a fresh network on n generated signals; returns (net, losses)."""
    t, d_t, clean, cfp = training_set(model, n, seed, stream=stream)
    return train_on(clean, d_t, model, epochs, batch, lr, snr, seed, stream)


def train_on(clean, d_t, model="NODDI_Bingham", epochs=20, batch=2048, lr=2e-3, snr=(10.0, 50.0), seed=1234,
             stream=None):
    """Train a fresh network on a generated training set (training_set's
    clean signals and truths); returns (net, losses)."""
    net = capi.Mlp(clean.shape[1], head(model), width=WIDTH, depth=DEPTH, seed=seed)
    losses = net.train(clean, None, d_t, snr=snr, epochs=epochs, batch=batch, lr=lr, seed=seed + 1, stream=stream)
    return net, losses


def save(net, path):
    np.save(path, net.get_weights(), allow_pickle=False)


def load(model, path):
    cfp = synth.model_acquisition(model)
    net = capi.Mlp(cfp.shape[0], head(model), width=WIDTH, depth=DEPTH)
    net.set_weights(np.load(path, allow_pickle=False))
    return net
