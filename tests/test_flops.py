"""FLOP model (cudimot_amd/flops.py over the operation-counting build of the
plug-ins): every shipped plug-in evaluates through each of its cache hooks,
with the engine's measurement cache handed in where the plug-in declares
one, and gives finite counts."""
import numpy as np
import pytest


@pytest.mark.parametrize("model", ["Ball_1_Stick", "Ball_2_Sticks", "Ball_3_Sticks", "DiffusionTensor",
                                   "NODDI_Watson", "NODDI_Bingham", "NODDI_Bingham_2Fans", "toymodel"])
def test_per_eval_counts_every_plugin(model):
    from cudimot_amd import flops, synth
    cfp = synth.model_acquisition(model)
    t = synth.truth_for(model, 6, 2)
    ev = flops.per_eval(model, t, cfp, np.full((6, 1), 1000.0))
    vals = [v for v in ev.values() if isinstance(v, (float, np.floating))]
    assert all(np.isfinite(v) and v >= 0 for v in vals), ev
    assert ev["fwd32"] + ev["fwd64"] > 0


def test_bingham_mcmc_flops_use_the_per_measurement_cache():
    """The intra-axonal cache is recomputed only in steps of the parameters
    it depends on (kappa, beta, th, ph, psi)."""
    from cudimot_amd import flops, synth
    cfp = synth.model_acquisition("NODDI_Bingham")
    ev = flops.per_eval("NODDI_Bingham", synth.truth_for("NODDI_Bingham", 6, 2), cfp, np.full((6, 1), 1000.0))
    assert ev["mvc"] and ev["mvc_dep"] == 0x7C
    assert ev["mv64"] > 0 and ev["pmv64"] == 0
