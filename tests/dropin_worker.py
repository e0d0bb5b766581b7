"""Runs the reference's plug-ins, compiled unmodified into the engine library
oracle/_ref/dropin (CDM_LIBRARY points there), on the GPU: predicted signal,
GridSearch on the reference pipeline's grid and LM, for the seeded volumes of
tests/test_dropin.py.  Writes the results to the .npz named on the command
line.  A separate process, because a process loads one engine library."""
import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def main(out_path, inputs_path):
    import torch
    from cudimot_amd import capi
    assert os.environ.get("CDM_LIBRARY"), "CDM_LIBRARY must name the drop-in library"
    assert torch.cuda.is_available()
    inp = np.load(inputs_path)
    out = {"models": np.array(capi.list_models())}
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    for name in ("Ball_1_Stick", "NODDI_Watson"):
        m = capi.Model(name, os.path.join(REPO, "cudimot_amd", "mymodels", name, "modelpriors"))
        cfp, fx, meas, truth, p0 = (inp[name + "_" + k] for k in ("cfp", "fixp", "meas", "truth", "start"))
        n, nm = meas.shape
        d_fx = dev(fx) if m.fixp_tsize else None
        sig = torch.zeros((n, nm), device="cuda")
        m.predicted_signal(dev(truth.reshape(n, -1, 1)), dev(cfp), d_fx, sig)
        m.load_grid(str(inp[name + "_grid"]), is_text=True)
        if name == "NODDI_Watson":
            m.set_fixed("3,4")
        kept = torch.full((n,), -7, dtype=torch.int32, device="cuda")
        pg = dev(p0)
        m.gridsearch(dev(meas), pg, dev(cfp), d_fx, kept=kept)
        lm = capi.Model(name, os.path.join(REPO, "cudimot_amd", "mymodels", name, "modelpriors"))
        pl = dev(inp[name + "_lmstart"])
        lm.levmar(dev(meas), pl, dev(cfp), d_fx, max_iters=100)
        capi.sync()
        out[name + "_signal"] = sig.cpu().numpy()
        out[name + "_kept"] = kept.cpu().numpy()
        out[name + "_gs"] = pg.cpu().numpy()
        out[name + "_lm"] = pl.cpu().numpy()
        out[name + "_sizes"] = np.array([m.nparams, m.cfp_tsize, m.fixp_tsize])
    np.savez(out_path, **out)
    print("dropin worker ok", flush=True)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
