"""Short workload for hardware-counter passes (rocprofv3 --pmc).

LM (10 iterations) and MCMC (100 burn-in + 100 recorded jumps, a sample
every 10) on 131,072 NODDI-Bingham voxels with 99 measurements at SNR 25,
then a 512 MiB device-to-device copy whose FETCH_SIZE / WRITE_SIZE readings
calibrate the traffic counters (its true traffic is 512 MiB each way).

    python3 cudimot_amd/profile_workload.py
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
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
