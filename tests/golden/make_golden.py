"""Write the golden fixtures under tests/golden/ (run in a checkout that has
the reference, after ``python oracle/build_oracle.py`` built oracle/_ref).

The fixtures come from oracle/_ref/libref_models.so: the reference's own
plug-in headers (mymodels/<name>/modelfunctions.h and helpers, compiled
unmodified from /root/reference) driven by the CPU restatement of the
reference's fitting loops.  They hold inputs and outputs only:
  fwd_<model>.npz   params, cfp, fixp -> signal, derivatives
  fit_<model>.npz   a small noisy volume -> GridSearch kept index and
                    parameters, Levenberg-Marquardt parameters, iterations
                    and cost, a short MCMC sample set
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from oracle import bind  # noqa: E402
from tests import helpers as H  # noqa: E402

MODELS = ["Ball_1_Stick", "Ball_2_Sticks", "DiffusionTensor", "toymodel", "NODDI_Watson", "NODDI_Bingham",
          "NODDI_Bingham_2Fans"]

# Reference modelpriors values (bounds/priors as parsed; see test_spec.py).
GRIDS = {
    "Ball_1_Stick": ([1, 2, 3, 4], [[0.0005, 0.001, 0.002, 0.003], [0.1, 0.4, 0.7], [0.3, 1.2, 2.1], [-2, 0, 2]]),
    "Ball_2_Sticks": ([2, 5], [[0.2, 0.4, 0.6], [0.0, 0.1, 0.2]]),
    "DiffusionTensor": ([1, 4, 6], [[0.0005, 0.0015, 0.0025]] * 3),
    "toymodel": ([0, 1], [[0.5, 1, 1.5, 2], [0.25, 0.5, 1, 1.5, 2]]),
    "NODDI_Watson": ([0, 1, 2], [[0.0, 0.1, 0.2, 0.3], [0.2, 0.5, 0.8], [1, 2, 4, 8]]),
    "NODDI_Bingham": ([0, 1, 2, 3], [[0.0, 0.2], [0.5, 0.8], [2, 8, 20], [0.5, 2, 9]]),
    "NODDI_Bingham_2Fans": ([0, 1], [[0.0, 0.2], [0.2, 0.4]]),
}


def cartesian(vals):
    grids = np.meshgrid(*vals, indexing="ij")
    return np.stack([g.ravel() for g in grids], 1)


def spec_for(model):
    sys.path.insert(0, REPO)
    from cudimot_amd import capi
    s = capi.Model(model).spec()
    return {k: np.asarray(v) for k, v in s.items() if k != "init_provided"}


def start_params(model, spec, t, noisy, cfp):
    p = np.tile(spec["p_init"], (len(t), 1))
    if model.startswith("Ball") or model == "DiffusionTensor":
        p[:, 0] = noisy[:, cfp[:, 3] == 0].mean(1)
    if model.startswith("NODDI"):
        # orientations from the truth (as a DTI initialisation would give)
        if model == "NODDI_Watson":
            p[:, 3:5] = t[:, 3:5]
        elif model == "NODDI_Bingham":
            p[:, 4:7] = t[:, 4:7]
        else:
            p[:, 5:8], p[:, 11:14] = t[:, 5:8], t[:, 11:14]
    return p


def main():
    if not bind.ref_available():
        raise SystemExit("oracle/_ref not built (needs the reference checkout)")
    rng = np.random.default_rng(20240601)
    for model in MODELS:
        r = bind.OracleModel(model, ref=True)
        n = 24
        cfp = H.ACQ[H.MODEL_ACQ[model]]()
        t = H.truth(model, n, seed=101)
        fx = H.fixp_for(model, n)
        sig = r.signal(t, cfp, fx)
        der = r.derivs(t, cfp, fx)
        np.savez_compressed(os.path.join(HERE, "fwd_%s.npz" % model), params=t, cfp=cfp, fixp=fx, signal=sig,
                            derivs=der)
        # fits on a small noisy volume
        nf = 24 if model.startswith("NODDI_Bingham") else 48
        cfp_f, tf, fxf, noisy = H.volume(r, model, nf, seed=202)
        spec = spec_for(model)
        p0 = start_params(model, spec, tf, noisy, cfp_f)
        gp, vals = GRIDS[model]
        grid = cartesian(vals)
        p_gs, kept = r.gridsearch(spec, noisy, p0, cfp_f, fxf, np.asarray(gp), grid)
        p_lm, it_lm, c_lm = r.levmar(spec, noisy, p_gs, cfp_f, fxf, max_iters=100)
        p_mc, samples, _ = r.mcmc(spec, noisy, p_lm, cfp_f, fxf, nburn=20, njumps=40, sampleevery=4, seed=77)
        np.savez_compressed(os.path.join(HERE, "fit_%s.npz" % model), cfp=cfp_f, fixp=fxf, meas=noisy, start=p0,
                            grid_params=np.asarray(gp), grid=grid, gs_params=p_gs, gs_kept=kept, lm_params=p_lm,
                            lm_iters=it_lm, lm_cost=c_lm, mcmc_samples=samples)
        print(model, "ok")


if __name__ == "__main__":
    main()
