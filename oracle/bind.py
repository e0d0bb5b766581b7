"""ORACLE ctypes bindings (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg import this module, and only as the checker.  It wraps
``oracle/_build/liboracle.so`` (FP64 restatement of the reference's fitting
loops over this project's plug-ins) and, when built, ``oracle/_ref/
libref_models.so`` (the same loops over the reference's own plug-in headers).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_models.so")

_libs = {}

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)


def _check_fresh(path):
    """liboracle.so must come from the sources in this tree: its build stamp
    holds their digest."""
    import sys
    sys.path.insert(0, HERE)
    try:
        import build_oracle
    finally:
        sys.path.pop(0)
    stamp = path + ".sha"
    have = open(stamp).read().strip() if os.path.exists(stamp) else None
    if have != build_oracle.sources_digest():
        raise RuntimeError("oracle library %s was built from other sources than this tree's "
                           "(run python oracle/build_oracle.py)" % path)


def _load(path):
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError("oracle library not built: %s (run python oracle/build_oracle.py)" % path)
        if path == LIB_PATH:
            _check_fresh(path)
        _libs[path] = ctypes.CDLL(path)
    return _libs[path]


def ref_available():
    return os.path.exists(REF_PATH)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_F)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


def philox4x32_10(ctr, key):
    lib = _load(LIB_PATH)
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (ctypes.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return [int(x) for x in o]


def omp_threads():
    """Threads the oracle's OpenMP loops run on (the team actually formed)."""
    return int(_load(LIB_PATH).orc_omp_threads())


def draws(seed, voxel, jump, slot):
    lib = _load(LIB_PATH)
    z = ctypes.c_double()
    u = ctypes.c_double()
    lib.orc_draws(ctypes.c_ulonglong(seed), ctypes.c_ulonglong(voxel), ctypes.c_uint32(jump),
                  ctypes.c_uint32(slot), ctypes.byref(z), ctypes.byref(u))
    return z.value, u.value


def _seconds_ptr(seconds, nvox):
    """Per-voxel wall-clock output of a fit (a float64 array of nvox, or None)."""
    if seconds is None:
        return None
    assert seconds.dtype == np.float64 and seconds.flags.c_contiguous and seconds.size >= nvox
    return seconds.ctypes.data_as(_D)


class OracleModel:
    """FP64 restatement of the reference's fitting loops for one plug-in.

    ``ref=True`` selects the instantiation over the reference's own plug-in
    headers (only where oracle/_ref was built)."""

    def __init__(self, name, ref=False):
        self.name = name
        self.lib = _load(REF_PATH if ref else LIB_PATH)
        self.prefix = ("ref_%s_" if ref else "orc_%s_") % name
        n, c, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._fn("info")(ctypes.byref(n), ctypes.byref(c), ctypes.byref(f))
        self.nparams, self.cfp_tsize, self.fixp_tsize = n.value, c.value, f.value

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _fixp(self, fixp, nvox):
        if fixp is None or self.fixp_tsize == 0:
            return np.zeros((nvox, max(self.fixp_tsize, 1)))
        return fixp

    def signal(self, params, cfp, fixp=None):
        params = np.atleast_2d(params)
        nvox, nmeas = params.shape[0], cfp.shape[0]
        p, pp = _d(params)
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        out = np.zeros((nvox, nmeas))
        self._fn("signal")(nvox, nmeas, pp, cp, xp, out.ctypes.data_as(_D))
        return out

    def derivs(self, params, cfp, fixp=None):
        params = np.atleast_2d(params)
        nvox, nmeas = params.shape[0], cfp.shape[0]
        p, pp = _d(params)
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        out = np.zeros((nvox, nmeas, self.nparams))
        self._fn("derivs")(nvox, nmeas, pp, cp, xp, out.ctypes.data_as(_D))
        return out

    def fix_constraints_lm(self, params):
        p = np.array(np.atleast_2d(params), dtype=np.float64)
        self._fn("fix_constraints_lm")(p.shape[0], p.ctypes.data_as(_D))
        return p

    def constraints_mcmc(self, params):
        p, pp = _d(np.atleast_2d(params))
        ok = np.zeros(p.shape[0], dtype=np.int32)
        self._fn("constraints_mcmc")(p.shape[0], pp, ok.ctypes.data_as(_I))
        return ok.astype(bool)

    def gridsearch(self, spec, meas, params, cfp, fixp, grid_params, grid, seconds=None):
        nvox, nmeas = meas.shape
        m, mp = _d(meas)
        p = np.array(params, dtype=np.float64, order="C")
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        bt, btp = _i(spec["bound_types"])
        lo, lop = _f(spec["bmin"])
        hi, hip = _f(spec["bmax"])
        gp, gpp = _i(grid_params)
        g, gptr = _d(grid)
        kept = np.zeros(nvox, dtype=np.int32)
        ncombs = g.shape[0] if len(gp) else 0
        self._fn("gridsearch")(btp, lop, hip, nvox, nmeas, mp, p.ctypes.data_as(_D), cp, xp, len(gp), gpp,
                               ncombs, gptr, kept.ctypes.data_as(_I), _seconds_ptr(seconds, nvox))
        return p, kept

    def levmar(self, spec, meas, params, cfp, fixp=None, max_iters=200, marquardt=True, seconds=None):
        nvox, nmeas = meas.shape
        m, mp = _d(meas)
        p = np.array(params, dtype=np.float64, order="C")
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        bt, btp = _i(spec["bound_types"])
        lo, lop = _f(spec["bmin"])
        hi, hip = _f(spec["bmax"])
        fx, fxp = _i(spec["fixed"])
        iters = np.zeros(nvox, dtype=np.int32)
        cost = np.zeros(nvox)
        self._fn("levmar")(btp, lop, hip, fxp, nvox, nmeas, mp, p.ctypes.data_as(_D), cp, xp, int(max_iters),
                           int(bool(marquardt)), iters.ctypes.data_as(_I), cost.ctypes.data_as(_D),
                           _seconds_ptr(seconds, nvox))
        return p, iters, cost

    def mcmc(self, spec, meas, params, cfp, fixp=None, nburn=1000, njumps=1250, sampleevery=25,
             update_prop=True, rician=False, seed=8219, voxel_offset=0, seconds=None):
        nvox, nmeas = meas.shape
        m, mp = _d(meas)
        p = np.array(params, dtype=np.float64, order="C")
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        arrs = [_i(spec["bound_types"]), _f(spec["bmin"]), _f(spec["bmax"]), _i(spec["prior_types"]),
                _f(spec["pa"]), _f(spec["pb"]), _i(spec["fixed"])]
        ns = njumps // sampleevery
        samples = np.zeros((nvox, self.nparams, ns))
        tau = np.zeros((nvox, ns))
        self._fn("mcmc")(*[a[1] for a in arrs], nvox, nmeas, mp, p.ctypes.data_as(_D), cp, xp, int(nburn),
                         int(njumps), int(sampleevery), int(bool(update_prop)), int(bool(rician)),
                         ctypes.c_ulonglong(seed), ctypes.c_longlong(voxel_offset), samples.ctypes.data_as(_D),
                         tau.ctypes.data_as(_D), _seconds_ptr(seconds, nvox))
        return p, samples, (tau if rician else None)

    def predicted(self, samples, cfp, fixp=None):
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        nvox, _, ns = samples.shape
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        out = np.zeros((nvox, cfp.shape[0]))
        self._fn("predicted")(nvox, cfp.shape[0], ns, samples.ctypes.data_as(_D), cp, xp, out.ctypes.data_as(_D))
        return out

    def bic_aic(self, meas, samples, cfp, fixp=None, rician=False, tau_samples=None):
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        nvox, _, ns = samples.shape
        m, mp = _d(meas)
        c, cp = _d(cfp)
        x, xp = _d(self._fixp(fixp, nvox))
        t = None
        if rician:
            t, tptr = _d(tau_samples)
        bic = np.zeros(nvox)
        aic = np.zeros(nvox)
        self._fn("bic_aic")(nvox, meas.shape[1], ns, int(bool(rician)), mp, samples.ctypes.data_as(_D),
                            tptr if rician else None, cp, xp, bic.ctypes.data_as(_D), aic.ctypes.data_as(_D))
        return bic, aic
