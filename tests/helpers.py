"""Shared fixtures of the parity tests: synthetic volumes per model and
device-buffer plumbing.  The oracle (oracle/bind.py) is the checker."""
import numpy as np

from cudimot_amd import synth

# Acquisitions of the measured workloads.
ACQ = {
    "M64": lambda: synth.acquisition(4, [(1000, 30), (2000, 30)], seed=11),
    "M99": lambda: synth.acquisition(9, [(700, 30), (2000, 60)], seed=12),
    "M280": lambda: synth.acquisition(10, [(1000, 90), (2000, 90), (3000, 90)], seed=13),
    "M20": lambda: np.linspace(0.0, 3.0, 20).reshape(-1, 1),  # toymodel x values
}

MODEL_ACQ = {"Ball_1_Stick": "M64", "Ball_2_Sticks": "M280", "Ball_3_Sticks": "M280", "DiffusionTensor": "M64",
             "toymodel": "M20", "NODDI_Watson": "M99", "NODDI_Watson_refnum": "M99", "NODDI_Bingham": "M99",
             "NODDI_Bingham_2Fans": "M99"}


def truth(model, n, seed):
    if model == "Ball_1_Stick":
        return synth.truth_ball_1_stick(n, seed)
    if model == "Ball_2_Sticks":
        return synth.truth_ball_k_sticks(n, 2, seed)
    if model == "Ball_3_Sticks":
        return synth.truth_ball_k_sticks(n, 3, seed)
    if model == "DiffusionTensor":
        return synth.truth_dti(n, seed)
    if model == "toymodel":
        return synth.truth_toymodel(n, seed)
    if model in ("NODDI_Watson", "NODDI_Watson_refnum"):
        return synth.truth_noddi_watson(n, seed)
    if model == "NODDI_Bingham":
        return synth.truth_noddi_bingham(n, seed)
    if model == "NODDI_Bingham_2Fans":
        t1 = synth.truth_noddi_bingham(n, seed)
        t2 = synth.truth_noddi_bingham(n, seed + 1)
        ff = np.random.default_rng(seed + 2).uniform(0.1, 0.5, n)
        return np.concatenate([t1[:, :1], ff[:, None], t1[:, 1:], t2[:, 1:]], 1)
    raise KeyError(model)


def fixp_for(model, n, s0=1000.0):
    if model.startswith("NODDI"):
        return np.full((n, 1), s0)
    return np.zeros((n, 1))


def volume(oracle_model, model, n, seed, snr=30.0):
    """(cfp, truth, fixp, noisy signal) from the oracle's FP64 forward model."""
    cfp = ACQ[MODEL_ACQ[model]]()
    t = truth(model, n, seed)
    fx = fixp_for(model, n)
    clean = oracle_model.signal(t, cfp, fx)
    s0 = fx[:, 0] if model.startswith("NODDI") else (t[:, 0] if model != "toymodel" else t[:, 0])
    noisy = synth.rician(clean, s0, snr, np.random.default_rng(seed + 100))
    return cfp, t, fx, noisy


def dev(a):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def host(t):
    return t.detach().cpu().numpy().astype(np.float64)
