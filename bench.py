"""Benchmark of the fitting path: NODDI-Bingham Levenberg-Marquardt + MCMC.

One step = LM (--iterLevMar 200) followed by MCMC (reference defaults:
1000 burn-in jumps, 1250 jumps, a sample every 25) over one synthetic
NODDI-Bingham volume of 500,000 voxels (9 b=0 + 30 directions at b=700 +
60 at b=2000, 99 measurements; kappa1 U[1,64], kappa2/kappa1 U[0,1],
psi U[0,pi), fiso U[0,0.3], fintra U[0.1,0.9], Rician noise at SNR 25),
already resident in HBM.  LM starts from the plug-in's P_init with the
fibre frame (th, ph, psi) from the ground truth plus 0.1 rad noise, as a
tensor-fit initialisation would give it.

Prints one JSON line: voxels/s (whole job), with
  roofline      the MCMC kernel (dominant): counted FP32/FP64 flops per
                launch over its average HIP-event duration, as a fraction of
                MI355X vector peaks (FP64 78.6, FP32 157.3 TFLOP/s):
                frac = F64/(78.6e12 t) + F32/(157.3e12 t); achieved is the
                FP64-equivalent rate
  cpu_baseline  the FP64 CPU oracle (OpenMP) on a bounded voxel sample
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_F64 = 78.6e12
PEAK_F32 = 157.3e12


def make_volume(m, n, seed, dev, torch):
    from cudimot_amd import synth
    cfp = synth.model_acquisition("NODDI_Bingham")
    # inside the model's bounds and constraint (beta <= 32, kappa - beta > 0.05)
    t = synth.truth_for("NODDI_Bingham", n, seed)
    d_cfp = dev(cfp)
    d_fx = torch.full((n, 1), 1000.0, device="cuda")
    clean = torch.empty((n, cfp.shape[0]), device="cuda")
    m.predicted_signal(dev(t.reshape(n, 7, 1)), d_cfp, d_fx, clean)
    g = torch.Generator(device="cuda").manual_seed(seed)
    sigma = 1000.0 / 25.0
    n1 = torch.randn(clean.shape, device="cuda", generator=g) * sigma
    n2 = torch.randn(clean.shape, device="cuda", generator=g) * sigma
    meas = torch.sqrt((clean + n1) ** 2 + n2 ** 2).contiguous()
    spec = m.spec()
    p0 = np.tile(np.asarray(spec["p_init"], dtype=np.float64), (n, 1))
    p0[:, 4:7] = t[:, 4:7] + 0.1 * np.random.default_rng(seed + 1).normal(size=(n, 3))
    return cfp, t, d_cfp, d_fx, meas, dev(p0)


def cpu_baseline(args, cfp, meas_np, p0_np, nsample):
    """FP64 CPU oracle (OpenMP, one voxel per iteration) on a voxel sample."""
    from oracle import bind
    from cudimot_amd import capi
    o = bind.OracleModel("NODDI_Bingham")
    spec = {k: np.asarray(v) for k, v in capi.Model("NODDI_Bingham").spec().items() if k != "init_provided"}
    fx = np.full((nsample, 1), 1000.0)
    t0 = time.time()
    p, _, _ = o.levmar(spec, meas_np, p0_np, cfp, fx, max_iters=args.lm_iters)
    o.mcmc(spec, meas_np, p, cfp, fx, nburn=args.burnin, njumps=args.jumps, sampleevery=args.sampleevery,
           seed=args.seed)
    dt = time.time() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": nsample / dt, "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": "%d voxels of the same volume, full LM (200 it.) + MCMC (%d+%d jumps) in %.1f s" %
                      (nsample, args.burnin, args.jumps, dt)}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--voxels", type=int, default=500000)
    ap.add_argument("--warmup-voxels", type=int, default=8192, help="voxels of each (untimed) warm-up step")
    ap.add_argument("--lm-iters", type=int, default=200)
    ap.add_argument("--burnin", type=int, default=1000)
    ap.add_argument("--jumps", type=int, default=1250)
    ap.add_argument("--sampleevery", type=int, default=25)
    ap.add_argument("--seed", type=int, default=8219)
    ap.add_argument("--lanes", type=int, default=0, help="voxel-group size for LM and MCMC (0: built-in choice)")
    ap.add_argument("--cpu-sample", type=int, default=32, help="voxels of the CPU-oracle baseline (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build fits one GPU's voxel shard and has no multi-GPU path "
              "(voxels shard with no exchange; see the design notes).", file=sys.stderr)
        sys.exit(2)

    import torch
    from cudimot_amd import capi, flops

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    m = capi.Model("NODDI_Bingham")
    if args.lanes:
        m.set_group(capi.LEVMAR, args.lanes)
        m.set_group(capi.MCMC, args.lanes)
    stream = torch.cuda.current_stream()
    P = m.nparams
    nsamp = args.jumps // args.sampleevery

    def step(meas, p0, d_cfp, d_fx, n, ev_mcmc=None):
        p = p0.clone()
        samples = torch.empty((n, P, nsamp), device="cuda")
        m.levmar(meas, p, d_cfp, d_fx, max_iters=args.lm_iters, stream=stream)
        if ev_mcmc:
            ev_mcmc[0].record(stream)
        m.mcmc(meas, p, d_cfp, d_fx, samples, None, nburn=args.burnin, njumps=args.jumps,
               sampleevery=args.sampleevery, seed=args.seed, stream=stream)
        if ev_mcmc:
            ev_mcmc[1].record(stream)
        return p, samples

    # warm-up on a small volume (module load, first launches)
    wc = make_volume(m, args.warmup_voxels, args.seed + 7, dev, torch)
    for _ in range(args.warmup):
        step(wc[4], wc[5], wc[2], wc[3], args.warmup_voxels)
    torch.cuda.synchronize()

    cfp, truth, d_cfp, d_fx, meas, p0 = make_volume(m, args.voxels, args.seed, dev, torch)
    torch.cuda.synchronize()
    print("bench: warm-up done, timing %d step(s) of %d voxels" % (args.steps, args.voxels), file=sys.stderr,
          flush=True)
    mcmc_ms = []
    t0 = time.time()
    for k in range(args.steps):
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        step(meas, p0, d_cfp, d_fx, args.voxels, ev)
        mcmc_ms.append(ev)
        if k + 1 < args.steps:
            torch.cuda.synchronize()
            print("bench: step %d done at %.1f s" % (k + 1, time.time() - t0), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.time() - t0
    mcmc_s = np.mean([a.elapsed_time(b) for a, b in mcmc_ms]) / 1000.0

    value = args.voxels * args.steps / elapsed
    # counted work of the MCMC call (burn-in + recording launches)
    sub = np.random.default_rng(0).choice(args.voxels, 256, replace=False)
    ev = flops.per_eval("NODDI_Bingham", truth[sub], cfp, np.full((256, 1), 1000.0))
    f32, f64 = flops.mcmc_flops(ev, P, cfp.shape[0], args.burnin + args.jumps)
    f32 *= args.voxels
    f64 *= args.voxels
    frac = f64 / (PEAK_F64 * mcmc_s) + f32 / (PEAK_F32 * mcmc_s)
    roofline = {"bound": "mfma", "pipe": "vector ALU (FP64 + FP32), no matrix cores", "kernel": "cdm_mcmc_kernel",
                "achieved": round(frac * PEAK_F64 / 1e12, 3), "peak": PEAK_F64 / 1e12, "unit": "TFLOP/s",
                "frac": round(frac, 4), "traffic": None,
                "flops_per_launch_f32": f32 / 2, "flops_per_launch_f64": f64 / 2,
                "avg_launch_ms": mcmc_s * 1000 / 2,
                "per_eval": {k: round(float(v), 2) for k, v in ev.items()}}

    out = {"metric": "voxels/s NODDI-Bingham LM+MCMC", "value": round(value, 2), "unit": "voxels/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed * 1000 / args.steps, 1),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "NODDI_Bingham LM(200)+MCMC(1000/1250/25), 99 meas, SNR 25",
                      "voxels": args.voxels, "lanes_lm": m.get_group(capi.LEVMAR, cfp.shape[0]),
                      "lanes_mcmc": m.get_group(capi.MCMC, cfp.shape[0])},
           "roofline": roofline}
    if args.cpu_sample > 0:
        ns = args.cpu_sample
        print("bench: GPU done (%.1f voxels/s); CPU oracle on %d voxels" % (value, ns), file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args, cfp, meas[:ns].cpu().numpy().astype(np.float64),
                                           p0[:ns].cpu().numpy().astype(np.float64), ns)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
