"""ctypes binding of the engine's C-ABI (``include/cudimot_amd.h``).

Host plumbing only: device buffers are torch tensors (or raw device
pointers); every call goes to ``lib/libcudimot.so``.  There is no CPU
fallback: if the library is missing or a call fails, an exception is raised.
"""
import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CDM_LIBRARY") or os.path.join(PKG, "lib", "libcudimot.so")

GRIDSEARCH, LEVMAR, MCMC, PREDICTED, BIC_AIC = range(5)

_lib = None


class CdmError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CdmError("engine library not built: %s (run python -m cudimot_amd.build)" % LIB_PATH)
        # PyTorch-ROCm wheels carry their own HIP/HSA runtime with the same
        # sonames as /opt/rocm's (libamdhip64.so.7, libhsa-runtime64.so.1).
        # Loading torch first makes this library bind to that one runtime, so
        # torch tensors, streams and events and this library's kernels share
        # queues and ordering.  Loaded the other way round, two runtimes
        # coexist in the process and the second one finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        vp, i, u64, i64, cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_longlong, ctypes.c_char_p
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        sig = {
            "cdm_version": (i, []),
            "cdm_last_error": (cp, []),
            "cdm_list_models": (i, [cp, i]),
            "cdm_model_open": (i, [cp, cp, ctypes.POINTER(vp)]),
            "cdm_model_close": (None, [vp]),
            "cdm_model_info": (i, [vp, ip, ip, ip]),
            "cdm_model_spec": (i, [vp, ip, dp, dp, ip, dp, dp, ip, dp, ip]),
            "cdm_model_set_fixed": (i, [vp, cp]),
            "cdm_model_load_grid": (i, [vp, cp, i, ip, ip]),
            "cdm_model_grid": (i, [vp, ip, dp]),
            "cdm_model_set_group": (i, [vp, i, i]),
            "cdm_model_get_group": (i, [vp, i, i, i]),
            "cdm_model_load_tuning": (i, [vp, cp]),
            "cdm_model_set_occupancy": (i, [vp, i, i]),
            "cdm_model_get_occupancy": (i, [vp, i, i]),
            "cdm_model_set_mcmc_blocks": (i, [vp, i, i]),
            "cdm_model_graph_stats": (i, [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]),
            "cdm_batch_graph_stats": (i, [ctypes.POINTER(i64), ctypes.POINTER(i64)]),
            "cdm_gridsearch": (i, [vp, i, i, vp, vp, vp, vp, vp, vp]),
            "cdm_levmar": (i, [vp, i, i, vp, vp, vp, vp, i, i, vp, vp, vp]),
            "cdm_mcmc": (i, [vp, i, i, vp, vp, vp, vp, i, i, i, i, i, u64, i64, vp, vp, vp]),
            "cdm_mcmc_noalloc": (i, [vp, i, i, vp, vp, vp, vp, i, i, i, i, i, u64, i64, vp, vp, vp]),
            "cdm_model_set_debug": (i, [vp, i, i]),
            "cdm_batch_set_fill": (i, [i]),
            "cdm_model_debug_trace": (i, [vp, i, vp, i, ip, vp]),
            "cdm_predicted_signal": (i, [vp, i, i, i, vp, vp, vp, vp, vp]),
            "cdm_bic_aic": (i, [vp, i, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp]),
            "cdm_rician_clean": (i, [i, i, vp]),
            "cdm_tensor_frame": (i, [i, vp, i, i, vp, vp]),
            "cdm_model_sizes": (i, [vp, ip, ip, ip, ip]),
            "cdm_sample_mean": (i, [i, i, i, vp, vp, vp]),
            "cdm_dyads": (i, [i, i, i, vp, i, i, i, vp, vp, vp]),
            "cdm_noddi_maps": (i, [i, i, i, vp, i, i, vp, vp]),
            "cdm_batch_units": (i, [cp, i]),
            "cdm_batch_levmar": (i, [i, vp, i, i, vp]),
            "cdm_batch_mcmc": (i, [i, vp, i, i, i, i, i, u64, i, i, vp]),
            "cdm_mlp_create": (i, [i, i, i, vp, i, u64, ctypes.POINTER(vp)]),
            "cdm_mlp_destroy": (None, [vp]),
            "cdm_mlp_set_train_graph": (i, [vp, i]),
            "cdm_mlp_info": (i, [vp, ip, ip, ip, ip, ctypes.POINTER(i64)]),
            "cdm_mlp_get_weights": (i, [vp, vp]),
            "cdm_mlp_set_weights": (i, [vp, vp]),
            "cdm_mlp_forward": (i, [vp, i, i, vp, vp, vp, vp]),
            "cdm_mlp_init_params": (i, [vp, i, i, vp, vp, i, vp, vp]),
            "cdm_mlp_loss_grad": (i, [vp, i, i, vp, vp, i, vp, vp, vp, vp]),
            "cdm_mlp_train": (i, [vp, i, i, vp, vp, i, vp, ctypes.c_float, ctypes.c_float, i, i, ctypes.c_float,
                                  u64, vp, vp]),
            "cdm_sync": (i, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise CdmError(lib().cdm_last_error().decode())
    return rc


def ptr(x):
    """Device pointer of a torch tensor (or an int pointer / None)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def list_models():
    n = lib().cdm_list_models(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().cdm_list_models(buf, n + 1)
    return [s for s in buf.value.decode().split(",") if s]


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Model:
    """A compiled model plug-in plus its parsed modelpriors, grid and
    --fixed list (the reference's Model<T>, Model.cu)."""

    def __init__(self, name, priors_file=None):
        self.name = name
        h = ctypes.c_void_p()
        check(lib().cdm_model_open(name.encode(), priors_file.encode() if priors_file else None, ctypes.byref(h)))
        self.h = h
        n, c, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().cdm_model_info(self.h, ctypes.byref(n), ctypes.byref(c), ctypes.byref(f)))
        self.nparams, self.cfp_tsize, self.fixp_tsize = n.value, c.value, f.value

    def close(self):
        if getattr(self, "h", None):
            lib().cdm_model_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spec(self):
        n = self.nparams
        I = ctypes.c_int * n
        D = ctypes.c_double * n
        bt, pt, fx = I(), I(), I()
        lo, hi, pa, pb, init = D(), D(), D(), D(), D()
        prov = ctypes.c_int()
        check(lib().cdm_model_spec(self.h, bt, lo, hi, pt, pa, pb, fx, init, ctypes.byref(prov)))
        return {"bound_types": list(bt), "bmin": list(lo), "bmax": list(hi), "prior_types": list(pt),
                "pa": list(pa), "pb": list(pb), "fixed": list(fx), "p_init": list(init),
                "init_provided": bool(prov.value)}

    def set_fixed(self, ids):
        s = ids if isinstance(ids, str) else ",".join(str(i) for i in ids)
        check(lib().cdm_model_set_fixed(self.h, s.encode()))

    def load_grid(self, src, is_text=False):
        nc, ng = ctypes.c_int(), ctypes.c_int()
        check(lib().cdm_model_load_grid(self.h, src.encode(), 1 if is_text else 0, ctypes.byref(nc),
                                        ctypes.byref(ng)))
        gp = (ctypes.c_int * max(ng.value, 1))()
        g = (ctypes.c_double * max(nc.value * ng.value, 1))()
        check(lib().cdm_model_grid(self.h, gp, g))
        flat = list(g)
        return list(gp)[:ng.value], [flat[k * ng.value:(k + 1) * ng.value] for k in range(nc.value)]

    def set_group(self, method, lanes):
        check(lib().cdm_model_set_group(self.h, method, lanes))

    def get_group(self, method, nmeas, nvox=0):
        r = lib().cdm_model_get_group(self.h, method, nmeas, int(nvox))
        if r < 0:
            raise CdmError(lib().cdm_last_error().decode())
        return r

    def set_occupancy(self, method, blocks):
        """Register budget of the LM / MCMC kernels (1, 2, 4 blocks per CU; 0: automatic)."""
        check(lib().cdm_model_set_occupancy(self.h, method, int(blocks)))

    def get_occupancy(self, method, nmeas):
        r = lib().cdm_model_get_occupancy(self.h, method, nmeas)
        if r < 0:
            raise CdmError(lib().cdm_last_error().decode())
        return r

    def load_tuning(self, path):
        """Tuning table (lines '<model> <nmeas> <method> <lanes>'); None clears it."""
        r = lib().cdm_model_load_tuning(self.h, path.encode() if path else None)
        if r < 0:
            raise CdmError(lib().cdm_last_error().decode())
        return r

    def set_mcmc_blocks(self, jumps_per_launch, use_graph=False):
        check(lib().cdm_model_set_mcmc_blocks(self.h, int(jumps_per_launch), 1 if use_graph else 0))

    def set_debug(self, voxel, max_records=4096):
        """--debug=<voxel>: GridSearch, LM and MCMC record that voxel's state
        (voxel < 0: off)."""
        check(lib().cdm_model_set_debug(self.h, int(voxel), int(max_records)))

    def debug_trace(self, method, stream=None):
        """The last call's records of `method` (GRIDSEARCH, LEVMAR, MCMC):
        array [n][4 + nparams] (index, three values, the parameters)."""
        import numpy as np
        n = ctypes.c_int()
        check(lib().cdm_model_debug_trace(self.h, method, None, 0, ctypes.byref(n), _stream(stream)))
        out = np.zeros((max(n.value, 0), 4 + self.nparams))
        if n.value:
            check(lib().cdm_model_debug_trace(self.h, method, out.ctypes.data_as(ctypes.c_void_p), n.value,
                                              ctypes.byref(n), _stream(stream)))
        return out

    def graph_stats(self):
        """(graphs recorded, graphs replayed from the handle's cache)."""
        c, r = ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().cdm_model_graph_stats(self.h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    # --- methods on device buffers (torch CUDA tensors, float32, contiguous)
    def gridsearch(self, meas, params, cfp, fixp=None, kept=None, stream=None):
        nvox, nmeas = meas.shape
        check(lib().cdm_gridsearch(self.h, nvox, nmeas, ptr(meas), ptr(params), ptr(cfp), ptr(fixp), ptr(kept),
                                   _stream(stream)))

    def levmar(self, meas, params, cfp, fixp=None, max_iters=200, marquardt=True, iters=None, cost=None,
               stream=None):
        nvox, nmeas = meas.shape
        check(lib().cdm_levmar(self.h, nvox, nmeas, ptr(meas), ptr(params), ptr(cfp), ptr(fixp), int(max_iters),
                               1 if marquardt else 0, ptr(iters), ptr(cost), _stream(stream)))

    def mcmc(self, meas, params, cfp, fixp, samples, tau_samples=None, nburn=1000, njumps=1250, sampleevery=25,
             update_proposal=True, rician=False, seed=8219, voxel_offset=0, stream=None):
        nvox, nmeas = meas.shape
        check(lib().cdm_mcmc(self.h, nvox, nmeas, ptr(meas), ptr(params), ptr(cfp), ptr(fixp), int(nburn),
                             int(njumps), int(sampleevery), 1 if update_proposal else 0, 1 if rician else 0,
                             int(seed), int(voxel_offset), ptr(samples), ptr(tau_samples), _stream(stream)))

    def mcmc_noalloc(self, meas, params, cfp, fixp, samples, tau_samples=None, nburn=1000, njumps=1250,
                     sampleevery=25, update_proposal=True, rician=False, seed=8219, voxel_offset=0, stream=None):
        """cdm_mcmc_noalloc: MCMC with the chain state in the library's static
        device scratch (no device allocation; a bounded voxel count)."""
        nvox, nmeas = meas.shape
        check(lib().cdm_mcmc_noalloc(self.h, nvox, nmeas, ptr(meas), ptr(params), ptr(cfp), ptr(fixp), int(nburn),
                                     int(njumps), int(sampleevery), 1 if update_proposal else 0, 1 if rician else 0,
                                     int(seed), int(voxel_offset), ptr(samples), ptr(tau_samples), _stream(stream)))

    def predicted_signal(self, samples, cfp, fixp, out, stream=None):
        nvox, _, ns = samples.shape
        check(lib().cdm_predicted_signal(self.h, nvox, out.shape[1], ns, ptr(samples), ptr(cfp), ptr(fixp),
                                         ptr(out), _stream(stream)))

    def bic_aic(self, meas, samples, cfp, fixp, bic, aic, rician=False, tau_samples=None, stream=None):
        nvox, nmeas = meas.shape
        ns = samples.shape[2]
        check(lib().cdm_bic_aic(self.h, nvox, nmeas, ns, 1 if rician else 0, ptr(meas), ptr(samples),
                                ptr(tau_samples), ptr(cfp), ptr(fixp), ptr(bic), ptr(aic), _stream(stream)))


class BatchItem(ctypes.Structure):
    """cdm_batch_item."""
    _fields_ = [("model", ctypes.c_void_p), ("nvox", ctypes.c_int), ("nmeas", ctypes.c_int),
                ("meas", ctypes.c_void_p), ("params", ctypes.c_void_p), ("cfp", ctypes.c_void_p),
                ("fixp", ctypes.c_void_p), ("iters", ctypes.c_void_p), ("cost", ctypes.c_void_p),
                ("samples", ctypes.c_void_p), ("tau_samples", ctypes.c_void_p), ("voxel_offset", ctypes.c_longlong)]


def batch_set_fill(blocks_per_cu):
    """Items filling at least blocks_per_cu blocks per CU run on their own
    kernels in the batched calls (default 8; 0: every item in the persistent
    launch)."""
    check(lib().cdm_batch_set_fill(int(blocks_per_cu)))


def batch_graph_stats():
    """(graphs recorded, graphs replayed) by cdm_batch_mcmc in this process."""
    c, r = ctypes.c_longlong(), ctypes.c_longlong()
    check(lib().cdm_batch_graph_stats(ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


def batch_units():
    n = lib().cdm_batch_units(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib().cdm_batch_units(buf, n + 1)
    out = {}
    for unit in buf.value.decode().split(";"):
        if unit:
            name, models = unit.split(":", 1)
            out[name] = models.split(",")
    return out


def _batch_items(items):
    """items: dicts with model (Model), meas, params, cfp, fixp (optional),
    iters, cost, samples, tau_samples (optional) and voxel_offset (optional)."""
    arr = (BatchItem * len(items))()
    for k, it in enumerate(items):
        meas = it["meas"]
        arr[k] = BatchItem(it["model"].h.value, meas.shape[0], meas.shape[1], ptr(meas), ptr(it["params"]),
                           ptr(it["cfp"]), ptr(it.get("fixp")), ptr(it.get("iters")), ptr(it.get("cost")),
                           ptr(it.get("samples")), ptr(it.get("tau_samples")), int(it.get("voxel_offset", 0)))
    return arr


def batch_levmar(items, max_iters=200, marquardt=True, stream=None):
    arr = _batch_items(items)
    check(lib().cdm_batch_levmar(len(items), ctypes.cast(arr, ctypes.c_void_p), int(max_iters),
                                 1 if marquardt else 0, _stream(stream)))


def batch_mcmc(items, nburn=1000, njumps=1250, sampleevery=25, update_proposal=True, seed=8219, jumps_per_launch=0,
               use_graph=False, rician=False, stream=None):
    arr = _batch_items(items)
    check(lib().cdm_batch_mcmc(len(items), ctypes.cast(arr, ctypes.c_void_p), int(nburn), int(njumps),
                               int(sampleevery), 1 if update_proposal else 0, 1 if rician else 0, int(seed),
                               int(jumps_per_launch), 1 if use_graph else 0, _stream(stream)))


def tensor_frame(dti_params, out, first=1, stream=None):
    """theta, phi, psi and eigenvalues [n][6] from tensor parameters [n][stride]."""
    n, stride = dti_params.shape
    check(lib().cdm_tensor_frame(n, ptr(dti_params), stride, first, ptr(out), _stream(stream)))


def sample_mean(samples, out, stream=None):
    """Mean over samples: samples [n][P][S] -> out [n][P] (fslmaths -Tmean)."""
    n, P, S = samples.shape
    check(lib().cdm_sample_mean(n, P, S, ptr(samples), ptr(out), _stream(stream)))


def dyads(samples, theta, phi, out, dispersion=None, psi=-1, stream=None):
    """Mean orientation [n][3] (and dispersion [n]) of the sampled directions
    (theta, phi), or of the Bingham fanning directions when psi >= 0."""
    n, P, S = samples.shape
    check(lib().cdm_dyads(n, P, S, ptr(samples), theta, phi, psi, ptr(out), ptr(dispersion), _stream(stream)))


def noddi_maps(samples, kappa, out, beta=-1, stream=None):
    """Watson (beta < 0): OD [n][1].  Bingham: [n][5] = mean k2, ODIp, ODIs,
    ODItot, DA."""
    n, P, S = samples.shape
    check(lib().cdm_noddi_maps(n, P, S, ptr(samples), kappa, beta, ptr(out), _stream(stream)))


def sync(stream=None):
    check(lib().cdm_sync(_stream(stream)))


def rician_clean(meas_np):
    """In-place --rician cleanup of a host float32 array [nvox][nmeas]."""
    import numpy as np
    assert meas_np.dtype == np.float32 and meas_np.flags.c_contiguous
    check(lib().cdm_rician_clean(meas_np.shape[0], meas_np.shape[1], meas_np.ctypes.data))


HEAD_SIGMOID, HEAD_SIGMOID_LOG, HEAD_SIGMOID_RATIO, HEAD_AXIS = range(4)


class HeadEntry(ctypes.Structure):
    """cdm_head_entry: how network outputs map to one model parameter."""
    _fields_ = [("kind", ctypes.c_int), ("param", ctypes.c_int), ("param2", ctypes.c_int), ("lo", ctypes.c_float),
                ("hi", ctypes.c_float), ("weight", ctypes.c_float), ("vmin", ctypes.c_float),
                ("vmax", ctypes.c_float)]


def head_entry(kind, param, param2=-1, lo=0.0, hi=1.0, weight=1.0, vmin=-3.0e38, vmax=3.0e38):
    return HeadEntry(kind, param, param2, lo, hi, weight, vmin, vmax)


class Mlp:
    """Learned initialiser network (cdm_mlp_*)."""

    def __init__(self, n_in, head, width=256, depth=3, seed=1):
        arr = (HeadEntry * len(head))(*head)
        h = ctypes.c_void_p()
        check(lib().cdm_mlp_create(int(n_in), int(width), int(depth), ctypes.cast(arr, ctypes.c_void_p), len(head),
                                   int(seed), ctypes.byref(h)))
        self.h = h
        self.head = list(head)
        a, b, c, d = (ctypes.c_int() for _ in range(4))
        nw = ctypes.c_longlong()
        check(lib().cdm_mlp_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d),
                                 ctypes.byref(nw)))
        self.n_in, self.width, self.depth, self.n_out, self.nweights = a.value, b.value, c.value, d.value, nw.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.cdm_mlp_destroy(self.h)
            self.h = None

    def layer_shapes(self):
        dims = [self.n_in] + [self.width] * self.depth + [self.n_out]
        return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]

    def get_weights(self):
        import numpy as np
        w = np.empty(self.nweights, dtype=np.float32)
        check(lib().cdm_mlp_get_weights(self.h, w.ctypes.data))
        return w

    def set_weights(self, w):
        import numpy as np
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == self.nweights
        check(lib().cdm_mlp_set_weights(self.h, w.ctypes.data))

    def forward(self, meas, scale, out, stream=None):
        n, nmeas = meas.shape
        check(lib().cdm_mlp_forward(self.h, n, nmeas, ptr(meas), ptr(scale), ptr(out), _stream(stream)))

    def init_params(self, meas, scale, params, stream=None):
        n, nmeas = meas.shape
        check(lib().cdm_mlp_init_params(self.h, n, nmeas, ptr(meas), ptr(scale), params.shape[1], ptr(params),
                                        _stream(stream)))

    def loss_grad(self, meas, scale, params_true, stream=None):
        import numpy as np
        n, nmeas = meas.shape
        loss = np.zeros(1, dtype=np.float32)
        grad = np.empty(self.nweights, dtype=np.float32)
        check(lib().cdm_mlp_loss_grad(self.h, n, nmeas, ptr(meas), ptr(scale), params_true.shape[1],
                                      ptr(params_true), loss.ctypes.data, grad.ctypes.data, _stream(stream)))
        return float(loss[0]), grad

    def set_train_graph(self, use_graph):
        """Training steps replayed from one captured HIP graph (default) or
        launched one by one; identical weights either way."""
        check(lib().cdm_mlp_set_train_graph(self.h, 1 if use_graph else 0))

    def train(self, clean, scale, params_true, snr=(10.0, 50.0), epochs=10, batch=4096, lr=1e-3, seed=1,
              stream=None):
        import numpy as np
        n, nmeas = clean.shape
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        check(lib().cdm_mlp_train(self.h, n, nmeas, ptr(clean), ptr(scale), params_true.shape[1], ptr(params_true),
                                  float(snr[0]), float(snr[1]), int(epochs), int(batch), float(lr), int(seed),
                                  losses.ctypes.data, _stream(stream)))
        return losses[:epochs]
