"""Short workloads for profiler runs.

Default (hardware-counter passes, rocprofv3 --pmc): LM (10 iterations) and
MCMC (100 burn-in + 100 recorded jumps, a sample every 10) on 131,072
NODDI-Bingham voxels with 99 measurements at SNR 25, then a 512 MiB
device-to-device copy whose FETCH_SIZE / WRITE_SIZE readings calibrate the
traffic counters (its true traffic is 512 MiB each way).

--train: the NODDI-Bingham initialiser's training as the benchmark runs it
(2,000,000 generated signals, 30 epochs, batches of 2,048), after a one-epoch
warm-up on 65,536 signals; prints the training time.

    python3 cudimot_amd/profile_workload.py [--train]
"""
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def train():
    import torch
    from cudimot_amd import learned
    learned.train("NODDI_Bingham", n=65536, epochs=1, seed=7)
    t, d_t, clean, cfp = learned.training_set("NODDI_Bingham", 2_000_000, 1234)
    torch.cuda.synchronize()
    t0 = time.time()
    net, losses = learned.train_on(clean, d_t, "NODDI_Bingham", epochs=30, seed=1234)
    torch.cuda.synchronize()
    print("trained 2000000 signals x 30 epochs in %.3f s (loss %.4f -> %.4f)" % (time.time() - t0, losses[0],
                                                                                 losses[-1]), flush=True)


def main():
    if "--train" in sys.argv[1:]:
        return train()
    import torch
    from cudimot_amd import capi, synth
    n = 131072
    cfp = synth.model_acquisition("NODDI_Bingham")
    t = synth.truth_for("NODDI_Bingham", n, 3)
    m = capi.Model("NODDI_Bingham")
    d_cfp = torch.as_tensor(cfp, dtype=torch.float32).cuda()
    s0 = torch.full((n, 1), 1000.0, device="cuda")
    clean = torch.empty((n, cfp.shape[0]), device="cuda")
    tt = torch.as_tensor(t, dtype=torch.float32).cuda()
    m.predicted_signal(tt.reshape(n, 7, 1), d_cfp, s0, clean)
    meas = torch.sqrt((clean + 40 * torch.randn_like(clean)) ** 2 + (40 * torch.randn_like(clean)) ** 2).contiguous()
    p = tt.clone()
    p[:, 2] *= 1.3
    m.levmar(meas, p, d_cfp, s0, max_iters=10)
    s = torch.empty((n, 7, 10), device="cuda")
    m.mcmc(meas, p, d_cfp, s0, s, None, nburn=100, njumps=100, sampleevery=10)
    a = torch.empty(128 * 1024 * 1024, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    print("done", flush=True)


if __name__ == "__main__":
    main()
