"""Algorithmic work model of the fitting methods (FLOPs by precision).

Invariants are counted once where they are invariant: a plug-in's voxel
cache (Voxel_Cache) once per parameter vector, the per-measurement cache
(Meas_Cache) once per launch (not counted), and the cached per-measurement
signal (Predicted_Signal_Cached) per evaluation.

Per forward evaluation F_fwd and per Partial_Derivatives call F_jac are
counted (not estimated) by running each plug-in on the CPU with an
operation-counting scalar (lib/libcdm_opcount.so, csrc/tools/opcount.h):
+, -, * = 1, / and sqrt = 1, each transcendental = 1, FMA = 2.  Method
totals per voxel:
  MCMC  = jumps * sum_params M * (F_meas + 3 FP32 + 1 FP64)  (residual,
          square, FP64 accumulation per measurement; F_meas is the cached
          per-measurement signal, plus the per-measurement cache when the
          proposed parameter is one it depends on; proposals and priors are
          O(P) per jump and not counted) + voxel caches of dependent proposals
  LM    = sum over iterations of M * (F_fwd + F_jac + P(P+1)/2 * 2 + 2P)
          for accepted steps plus M * F_fwd per cost evaluation
  GS    = combinations * M * (F_fwd + 3)
"""
import ctypes
import os

import numpy as np

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcdm_opcount.so")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("FLOP model not built: %s (python -m cudimot_amd.build)" % LIB)
        _lib = ctypes.CDLL(LIB)
    return _lib


def per_eval(model, params, cfp, fixp=None):
    """Average counts per (voxel, measurement): dict with fwd/jac FP32/FP64
    flops and transcendental calls."""
    params = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    cfp = np.ascontiguousarray(cfp, dtype=np.float64)
    n = params.shape[0]
    fixp = np.ascontiguousarray(fixp if fixp is not None else np.zeros((n, 1)), dtype=np.float64)
    out = np.zeros(20)
    D = ctypes.POINTER(ctypes.c_double)
    getattr(_load(), "cdm_opcount_" + model)(n, cfp.shape[0], params.ctypes.data_as(D), cfp.ctypes.data_as(D),
                                             fixp.ctypes.data_as(D), out.ctypes.data_as(D))
    return {"fwd32": out[0], "fwd64": out[1], "fwd_t32": out[2], "fwd_t64": out[3],
            "jac32": out[4], "jac64": out[5], "jac_t32": out[6], "jac_t64": out[7],
            "vc32": out[8], "vc64": out[9], "vc_dep": int(out[10]), "cached": bool(out[11]),
            "meas32": out[12], "meas64": out[13],
            "mv32": out[14], "mv64": out[15], "pmv32": out[16], "pmv64": out[17], "mvc_dep": int(out[18]),
            "mvc": bool(out[19])}


def mcmc_flops(ev, nparams, nmeas, jumps):
    """(FP32, FP64) flops of one voxel's MCMC over `jumps` jumps: every
    proposal evaluates the signal at all measurements, recomputes the voxel
    cache when the proposed parameter is one it depends on, and (with
    per-measurement caches) recomputes those only when it depends on them."""
    ndep = bin(ev["vc_dep"]).count("1") if ev["cached"] else 0
    f32 = jumps * ndep * ev["vc32"]
    f64 = jumps * ndep * ev["vc64"]
    for i in range(nparams):
        if ev.get("mvc"):
            dep = (ev["mvc_dep"] >> i) & 1
            m32 = ev["pmv32"] + (ev["mv32"] if dep else 0.0)
            m64 = ev["pmv64"] + (ev["mv64"] if dep else 0.0)
        else:
            m32, m64 = ev["meas32"], ev["meas64"]
        f32 += jumps * nmeas * (m32 + 3)
        f64 += jumps * nmeas * (m64 + 1)
    return f32, f64


def lm_flops(ev, nparams, nmeas, iters, jac_double=False):
    """(FP32, FP64) flops of one voxel's LM with `iters` accepted iterations
    (rejected steps cost one FP64 cost evaluation; counted as accepted ones
    in the cost term, a lower bound on the Jacobian term)."""
    tri = nparams * (nparams + 1) // 2
    jac32 = 0 if jac_double else ev["fwd32"] + ev["jac32"]
    jac64 = (ev["fwd32"] + ev["jac32"]) if jac_double else ev["fwd64"] + ev["jac64"]
    per_acc32 = nmeas * (jac32 + 2 * tri + 2 * nparams)
    per_acc64 = nmeas * jac64
    cost64 = nmeas * (ev["fwd32"] + ev["fwd64"] + 3)  # FP64 cost evaluation
    return iters * per_acc32, iters * per_acc64 + (iters + 1) * cost64
