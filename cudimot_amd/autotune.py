"""Autotuner of the launch shape: voxel-group size (lanes per voxel) and, for
LM and MCMC, register budget, per model, measurement count and method.

For every method it times the kernels at each compiled group size (1, 4,
16, 64 lanes) and register budget (1, 2, 4 resident blocks per CU) on a
synthetic volume large enough to fill the GPU, with a representative
workload (LM from perturbed starts with the default 200-iteration cap, on
four copies of the volume so that every resident group takes several voxels
from the work queue and the iteration-count spread and lane divergence show;
a few MCMC jumps), and keeps the fastest
(GridSearch keeps the built-in choice: its cost depends on the user's grid).
Results go to a text table, one line per choice:

    <model> <nmeas> <method> <lanes> <ms> <blocks>

which the C-ABI reads at model open (cdm_model_load_tuning) and this module
applies.  Run: python -m cudimot_amd.autotune [--models A,B] [--voxels N]
"""
import argparse
import os
import time

import numpy as np

from . import capi, synth

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.txt")
METHOD_NAMES = {capi.GRIDSEARCH: "gridsearch", capi.LEVMAR: "levmar", capi.MCMC: "mcmc",
                capi.PREDICTED: "predicted", capi.BIC_AIC: "bic_aic"}
LANES = (1, 4, 16, 64)
OCCUPANCY = (1, 2, 4)


def _volume(model, m, n, torch):
    cfp = synth.model_acquisition(model)
    t = synth.truth_for(model, n, seed=3)
    fx = torch.full((n, 1), 1000.0, device="cuda")
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    d_cfp = dev(cfp)
    clean = torch.empty((n, cfp.shape[0]), device="cuda")
    m.set_group(capi.PREDICTED, 16)
    m.predicted_signal(dev(t.reshape(n, -1, 1)), d_cfp, fx, clean)
    noise = 0.03 * clean.abs().max()
    meas = torch.sqrt((clean + noise * torch.randn_like(clean)) ** 2 + (noise * torch.randn_like(clean)) ** 2)
    return cfp, dev(t), d_cfp, fx, meas.contiguous()


def _lm_start(m, t, seed=4):
    """Starts near the truth (5% relative perturbation, inside the bounds):
    LM then runs a voxel-dependent number of iterations, as after GridSearch
    or the learned initialiser, so the timing sees the load imbalance."""
    spec = m.spec()
    rng = np.random.default_rng(seed)
    p = t * (1.0 + 0.05 * rng.normal(size=t.shape))
    for i, bt in enumerate(spec["bound_types"]):
        lo, hi = spec["bmin"][i], spec["bmax"][i]
        span = hi - lo if bt == 3 else 1.0
        if bt in (1, 3):
            p[:, i] = np.maximum(p[:, i], lo + 0.01 * span)
        if bt in (2, 3):
            p[:, i] = np.minimum(p[:, i], hi - 0.01 * span)
    return p


def tune_model(model, n=65536, reps=2):
    import torch
    m = capi.Model(model)
    cfp, t, d_cfp, fx, meas = _volume(model, m, n, torch)
    start = torch.as_tensor(np.ascontiguousarray(_lm_start(m, t.cpu().numpy()), dtype=np.float32)).cuda()
    # LM on four copies of the volume: a short or small LM timing favours
    # one lane per voxel, which loses twice over at full size and cap
    meas_lm, start_lm = meas.repeat(4, 1).contiguous(), start.repeat(4, 1).contiguous()
    fx_lm = fx.repeat(4, 1).contiguous() if fx is not None else None
    P = m.nparams
    samples = torch.empty((n, P, 2), device="cuda")
    out = torch.empty_like(meas)
    bic = torch.empty(n, device="cuda")
    aic = torch.empty(n, device="cuda")
    k = min(n, 4096)  # warm-up slice (first launch of each variant)
    runs = {
        capi.LEVMAR: (lambda: m.levmar(meas_lm, start_lm.clone(), d_cfp, fx_lm, max_iters=200),
                      lambda: m.levmar(meas[:k], start[:k].clone(), d_cfp, fx[:k], max_iters=5), 1),
        capi.MCMC: (lambda: m.mcmc(meas, t.clone(), d_cfp, fx, samples, None, nburn=16, njumps=16, sampleevery=8),
                    None, reps),
        capi.PREDICTED: (lambda: m.predicted_signal(samples, d_cfp, fx, out), None, reps),
        capi.BIC_AIC: (lambda: m.bic_aic(meas, samples, d_cfp, fx, bic, aic), None, reps),
    }
    res = {}
    for method, (fn, warm, nrep) in runs.items():
        best = None
        budgets = OCCUPANCY if method in (capi.LEVMAR, capi.MCMC) else (1,)
        for L in LANES:
            m.set_group(method, L)
            for W in budgets:
                if method in (capi.LEVMAR, capi.MCMC):
                    m.set_occupancy(method, W)
                (warm or fn)()
                torch.cuda.synchronize()
                t0 = time.time()
                for _ in range(nrep):
                    fn()
                torch.cuda.synchronize()
                ms = (time.time() - t0) * 1000 / nrep
                print("    %s %s lanes %d budget %d: %.3f ms" % (model, METHOD_NAMES[method], L, W, ms), flush=True)
                if best is None or ms < best[1]:
                    best = (L, ms, W)
        res[method] = best
        print("  %s %s: lanes %d, %.3f ms, budget %d" % (model, METHOD_NAMES[method], best[0], best[1], best[2]),
              flush=True)
    return cfp.shape[0], res


def write_table(rows, path=TABLE):
    keep = []
    if os.path.exists(path):
        new_keys = {(r[0], r[1], r[2]) for r in rows}
        for line in open(path):
            f = line.split()
            if len(f) >= 4 and not line.startswith("#") and (f[0], int(f[1]), f[2]) not in new_keys:
                keep.append(line.rstrip("\n"))
    with open(path, "w") as fh:
        fh.write("# model nmeas method lanes ms blocks  (written by python -m cudimot_amd.autotune)\n")
        for line in keep:
            fh.write(line + "\n")
        for model, nmeas, method, lanes, ms, occ in rows:
            fh.write("%s %d %s %d %.3f %d\n" % (model, nmeas, method, lanes, ms, occ))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--models", default="")
    ap.add_argument("--voxels", type=int, default=65536)
    ap.add_argument("--out", default=TABLE, help="table to write (default: the library's own)")
    a = ap.parse_args(argv)
    models = [x for x in a.models.split(",") if x] or capi.list_models()
    rows = []
    for model in models:
        nmeas, res = tune_model(model, a.voxels)
        for method, (L, ms, W) in res.items():
            rows.append((model, nmeas, METHOD_NAMES[method], L, ms, W))
        print(model, nmeas, {METHOD_NAMES[k]: v for k, v in res.items()}, flush=True)
    write_table(rows, a.out)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
