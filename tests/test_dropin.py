"""Plug-in drop-in: the reference's own model plug-ins (mymodels/Ball_1_Stick
and mymodels/NODDI_Watson: typedef double MyType, the *_gpu helpers,
NUMERICAL, their modelpriors) compiled unmodified by this project's build
(python -m cudimot_amd.build --models none --extra-model-dir DIR --out OUT)
into oracle/_ref/dropin, then

  * on the CPU: the library holds exactly those models, and cdm_model_open
    parses the reference's modelpriors files with the plug-ins' sizes;
  * on the GPU: predicted signal, GridSearch on the reference pipeline's
    grids and LM through the C-ABI agree with the oracle engine run over the
    same reference headers (oracle/_ref/libref_models.so).  Ball_1_Stick:
    signals within 1e-12 of the voxel's largest (plus the FP32 output
    rounding), the same GridSearch combination for every voxel, LM cost
    within 1e-6 for >= 97% of voxels.  NODDI_Watson: the reference computes
    erf in single precision (erff) and its Legendre-Gaussian recurrence
    amplifies one ulp of it into up to 1.4e-2 of the intra-cellular signal
    (kappa 30; 7.5e-4 at kappa 1), and the device and host math libraries'
    erff differ by an ulp for some arguments: signals within 2e-2, the
    GridSearch combination the same except where the oracle's own costs of
    the two combinations are within 1e-3, LM cost within 1e-3 for >= 90%.

The library is built by build() where the reference checkout exists (it is
git-ignored and travels with the tree); without it these tests skip.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import bind

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_MODELS = "/root/reference/mymodels"
DROPIN_LIB = os.path.join(REPO, "oracle", "_ref", "dropin", "lib", "libcudimot.so")


def _run_child(code, env_lib, timeout=300):
    env = dict(os.environ, CDM_LIBRARY=env_lib)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=REPO,
                       timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def test_reference_plugins_compile_unmodified():
    if not os.path.isdir(REF_MODELS):
        pytest.skip("the reference checkout is not present here")
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    try:
        import build_oracle
    finally:
        sys.path.pop(0)
    lib = build_oracle.build_dropin()
    assert lib == DROPIN_LIB and os.path.exists(lib)
    code = (
        "import json, numpy as np\n"
        "from cudimot_amd import capi\n"
        "out = {'models': capi.list_models()}\n"
        "for n in ('Ball_1_Stick', 'NODDI_Watson'):\n"
        "    m = capi.Model(n, '%s/' + n + '/modelpriors')\n"
        "    s = m.spec()\n"
        "    out[n] = dict(sizes=[m.nparams, m.cfp_tsize, m.fixp_tsize], bmin=list(map(float, s['bmin'])),\n"
        "                  bmax=list(map(float, s['bmax'])), bt=list(map(int, s['bound_types'])),\n"
        "                  pt=list(map(int, s['prior_types'])), p_init=list(map(float, s['p_init'])))\n"
        "print(json.dumps(out))\n" % REF_MODELS)
    got = json.loads(_run_child(code, lib).strip().splitlines()[-1])
    assert sorted(got["models"]) == ["Ball_1_Stick", "NODDI_Watson"]
    assert got["Ball_1_Stick"]["sizes"] == [5, 4, 0]
    assert got["NODDI_Watson"]["sizes"] == [5, 4, 1]
    # the reference's modelpriors: Ball_1_Stick P_init and bounds, Watson bounds and the sin prior on theta
    b = got["Ball_1_Stick"]
    assert b["p_init"] == [1.0, 0.001, 0.6, 1.5, 1.5]
    assert b["bt"][:3] == [1, 3, 3] and b["pt"][3] == 4
    assert b["bmin"][:3] == pytest.approx([1e-9, 1e-9, 0.01]) and b["bmax"][1:3] == pytest.approx([0.01, 1.0])
    w = got["NODDI_Watson"]
    assert w["bt"] == [3, 3, 3, 3, 3] and w["bmax"][:3] == [1.0, 1.0, 64.0] and w["pt"][3] == 4


def _inputs(path):
    from tests import helpers as H
    from tests.test_gpu_parity import GRIDS, _near_truth, _clip_noddi
    arrs = {}
    o1 = bind.OracleModel("Ball_1_Stick", ref=True)
    cfp, t, fx, noisy = H.volume(o1, "Ball_1_Stick", 256, seed=71)
    p0 = np.tile([1.0, 0.001, 0.6, 1.5, 1.5], (256, 1))
    p0[:, 0] = noisy[:, cfp[:, 3] == 0].mean(1)
    arrs.update(Ball_1_Stick_cfp=cfp, Ball_1_Stick_fixp=np.zeros((256, 1)), Ball_1_Stick_meas=noisy,
                Ball_1_Stick_truth=t, Ball_1_Stick_start=p0, Ball_1_Stick_lmstart=p0,
                Ball_1_Stick_grid=np.array(GRIDS["Ball_1_Stick"]))
    ow = bind.OracleModel("NODDI_Watson", ref=True)
    cfp, t, fx, noisy = H.volume(ow, "NODDI_Watson", 256, seed=72)
    tc = _clip_noddi("NODDI_Watson", t.copy())
    p0 = np.tile([0.4, 0.8, 5.0, 1.5, 1.5], (256, 1))
    p0[:, 3:5] = tc[:, 3:5]
    arrs.update(NODDI_Watson_cfp=cfp, NODDI_Watson_fixp=fx, NODDI_Watson_meas=noisy, NODDI_Watson_truth=tc,
                NODDI_Watson_start=p0, NODDI_Watson_lmstart=_near_truth("NODDI_Watson", t, np.random.default_rng(9)),
                NODDI_Watson_grid=np.array(GRIDS["NODDI_Watson"]))
    arrs = {k: (v.astype(np.float32).astype(np.float64) if v.dtype == np.float64 else v) for k, v in arrs.items()}
    np.savez(path, **arrs)
    return arrs


@pytest.mark.gpu
def test_reference_plugins_on_gpu_match_oracle(tmp_path):
    if not (os.path.exists(DROPIN_LIB) and bind.ref_available()):
        pytest.skip("oracle/_ref (the reference's plug-ins compiled from source) is not built here")
    from cudimot_amd import capi
    inp = _inputs(str(tmp_path / "in.npz"))
    out_path = str(tmp_path / "out.npz")
    _run_child("import sys; sys.path.insert(0, %r); from tests import dropin_worker as w; w.main(%r, %r)" %
               (REPO, out_path, str(tmp_path / "in.npz")), DROPIN_LIB, timeout=600)
    out = np.load(out_path)
    assert sorted(out["models"].tolist()) == ["Ball_1_Stick", "NODDI_Watson"]
    for name, sig_tol, tie, lm_tol, lm_frac in (("Ball_1_Stick", 1e-12, 0.0, 1e-6, 0.97),
                                                ("NODDI_Watson", 2e-2, 1e-3, 1e-3, 0.90)):
        o = bind.OracleModel(name, ref=True)
        g = lambda k: inp[name + "_" + k]
        fx = g("fixp")
        ref = o.signal(g("truth"), g("cfp"), fx)
        err = np.abs(out[name + "_signal"] - ref) / np.abs(ref).max(1, keepdims=True)
        # the GPU writes FP32 outputs: compare at float rounding plus the stated tolerance
        assert err.max() <= sig_tol + 6e-8, (name, err.max())
        fixed = "3,4" if name == "NODDI_Watson" else None
        mm = capi.Model(name)
        if fixed:
            mm.set_fixed(fixed)
        gp, grid = mm.load_grid(str(g("grid")), is_text=True)
        spec = {k: np.asarray(v) for k, v in mm.spec().items() if k != "init_provided"}
        _, kept_ref = o.gridsearch(spec, g("meas"), g("start"), g("cfp"), fx, np.asarray(gp), np.asarray(grid))
        mism = np.nonzero(out[name + "_kept"] != kept_ref)[0]

        def cost(v, k):
            q = g("start")[v].copy()
            q[np.asarray(gp)] = np.asarray(grid)[k]
            return ((o.signal(q[None], g("cfp"), fx[v:v + 1])[0] - g("meas")[v]) ** 2).sum()
        for v in mism:
            a, b = cost(v, out[name + "_kept"][v]), cost(v, kept_ref[v])
            assert abs(a - b) <= tie * b, (name, v, a, b)
        spec = {k: np.asarray(v) for k, v in capi.Model(name).spec().items() if k != "init_provided"}
        p_ref, _, _ = o.levmar(spec, g("meas"), g("lmstart"), g("cfp"), fx, max_iters=100)
        cost = lambda q: ((o.signal(q.astype(np.float32).astype(np.float64), g("cfp"), fx) - g("meas")) ** 2).sum(1)
        c_gpu, c_ref = cost(out[name + "_lm"]), cost(p_ref)  # both as returned, float-rounded
        agree = np.mean(np.abs(c_gpu - c_ref) / c_ref <= lm_tol)
        print(name, "drop-in: signal err max %.2e median %.2e, GridSearch mismatches %d, LM cost agreement %.3f" %
              (err.max(), np.median(err), mism.size, agree))
        assert agree >= lm_frac, (name, agree)
