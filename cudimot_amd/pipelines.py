"""Multi-step fitting recipes on the engine's C-ABI.

Restates the reference's pipeline scripts for the NODDI models
(mymodels/NODDI_Bingham/Pipeline_NODDI_Bingham.sh:177-232,
mymodels/NODDI_Watson/Pipeline_NODDI_Watson.sh:185-246) as calls on device
buffers, and adds the learned-initialiser pipeline that replaces their
GridSearch step:

  GridSearch + LM (Bingham), as the reference's script runs it:
    1. tensor fit (DiffusionTensor plug-in, LM) -> V1 -> theta, phi and
       V2 -> psi (Run_dtifit.sh, cart2spherical, initialise_Psi);
    2. GridSearch over fiso, fintra, kappa, beta with theta, phi fixed
       (--fixed=4,5, grid of Pipeline_NODDI_Bingham.sh:204-208), then LM with
       the same parameters fixed (the script's step 1 runs LM by default);
    3. LM of all parameters but fiso (--fixed=0).
  GridSearch + LM (Watson), Pipeline_NODDI_Watson.sh:185-246:
    1. tensor fit -> theta, phi of V1;
    2. GridSearch over fiso, fintra, kappa (grid :200-202) with theta, phi
       fixed (--fixed=3,4, --no_LevMar);
    3. LM of the fractions alone (--fixed=2,3,4);
    4. LM of all parameters.
  Learned + LM:
    1. the network predicts fiso, fintra, kappa, beta, theta, phi;
    2. psi from a 12-point scan (GridSearch over psi alone);
    3. LM of all parameters;
    4. fallback: voxels whose LM residual exceeds the noise bound (below) are
       refitted with GridSearch + LM, and the lower residual is kept.

Noise bound of the fallback: the SSE of a good fit is about (M - P) sigma^2.
sigma^2 is estimated from the spread of each voxel's b=0 measurements; the
volume's median of those estimates (divided by the median of chi^2_k / k,
k = n_b0 - 1) gives one robust sigma^2, and a voxel is refitted when its
SSE > factor (M - P) sigma^2 (factor 1.5: above the 99.9% point of
chi^2_{M-P} / (M - P) for M - P ~ 90).
"""
import numpy as np

from . import capi

BINGHAM_GRID = ("search[0]=(0.0,0.05,0.1,0.2,0.3)\n"
                "search[1]=(0.3,0.5,0.6,0.7,0.75,0.8,0.85,0.9,0.95,1.0)\n"
                "search[2]=(0.5,1,2,3,4,5,6,7,8,10,12,15,20,25,30,40,50)\n"
                "search[3]=(0.2,0.5,1.0,1.5,2.0,2.5,3.0,4.5,6.5,9,12,16,20)\n")
WATSON_GRID = ("search[0]=(0.0,0.1,0.2,0.3,0.4,0.5,0.6,0.7,0.8,0.9,1.0)\n"
               "search[1]=(0.0,0.1,0.2,0.3,0.4,0.5,0.6,0.7,0.8,0.9,1.0)\n"
               "search[2]=(1,2,3,4,5,6,7,8)\n")
PSI_SCAN = 12


def _torch():
    import torch
    return torch


def _dev(a):
    torch = _torch()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def b0_mask(cfp):
    return np.asarray(cfp)[:, 3] <= 50.0  # b <= 50 counts as b=0 (Pipeline_NODDI_Bingham.sh:155-160)


def tensor_frame(meas, cfp, iters=50, stream=None):
    """Tensor fit by LM (DiffusionTensor plug-in) started from S0 = mean b=0
    signal and D = 1e-3 I; returns the device frame [n][6] (theta, phi, psi,
    eigenvalues)."""
    torch = _torch()
    n = meas.shape[0]
    mask = torch.as_tensor(b0_mask(cfp), device=meas.device)
    s0 = meas[:, mask].mean(1)
    p = torch.zeros((n, 7), device=meas.device)
    p[:, 0] = s0
    p[:, 1] = p[:, 4] = p[:, 6] = 1e-3
    dti = capi.Model("DiffusionTensor")
    dti.levmar(meas, p, _dev(cfp), None, max_iters=iters, stream=stream)
    frame = torch.empty((n, 6), device=meas.device)
    capi.tensor_frame(p, frame, first=1, stream=stream)
    return frame


def _model(name, fixed=None, grid=None):
    m = capi.Model(name)
    if fixed:
        m.set_fixed(fixed)
    if grid:
        m.load_grid(grid, is_text=True)
    return m


def bingham_gridsearch_lm(meas, cfp, s0, lm_iters=200, grid=BINGHAM_GRID, frame=None, stream=None):
    """GridSearch + LM recipe; returns (params [n][7], cost [n]) on the device."""
    torch = _torch()
    n = meas.shape[0]
    d_cfp = _dev(cfp)
    if frame is None:
        frame = tensor_frame(meas, cfp, stream=stream)
    spec = capi.Model("NODDI_Bingham").spec()
    p = _dev(np.tile(np.asarray(spec["p_init"], dtype=np.float32), (n, 1)))
    p[:, 4:7] = frame[:, 0:3]
    step1 = _model("NODDI_Bingham", fixed="4,5", grid=grid)
    step1.gridsearch(meas, p, d_cfp, s0, stream=stream)
    step1.levmar(meas, p, d_cfp, s0, max_iters=lm_iters, stream=stream)
    step2 = _model("NODDI_Bingham", fixed="0")
    cost = torch.empty(n, device=meas.device)
    step2.levmar(meas, p, d_cfp, s0, max_iters=lm_iters, cost=cost, stream=stream)
    return p, cost


def watson_gridsearch_lm(meas, cfp, s0, lm_iters=200, grid=WATSON_GRID, frame=None, kept=None, stream=None,
                         model="NODDI_Watson"):
    """GridSearch + LM recipe of the reference's Watson pipeline; returns
    (params [n][5], cost [n]) on the device.  kept (device int [n],
    optional) receives the GridSearch combination per voxel.  model
    "NODDI_Watson_refnum" runs the recipe with the reference's own numerical
    forms (closed-form SH coefficients for 0.4 < kappa <= 3, single-precision
    erf in the Legendre-Gaussian integrals)."""
    torch = _torch()
    n = meas.shape[0]
    d_cfp = _dev(cfp)
    if frame is None:
        frame = tensor_frame(meas, cfp, stream=stream)
    spec = capi.Model(model).spec()
    p = _dev(np.tile(np.asarray(spec["p_init"], dtype=np.float32), (n, 1)))
    p[:, 3:5] = frame[:, 0:2]
    _model(model, fixed="3,4", grid=grid).gridsearch(meas, p, d_cfp, s0, kept=kept, stream=stream)
    _model(model, fixed="2,3,4").levmar(meas, p, d_cfp, s0, max_iters=lm_iters, stream=stream)
    cost = torch.empty(n, device=meas.device)
    capi.Model(model).levmar(meas, p, d_cfp, s0, max_iters=lm_iters, cost=cost, stream=stream)
    return p, cost


def noise_sigma2(meas, cfp):
    """Volume-wide sigma^2 from the spread of the b=0 measurements."""
    from scipy.stats import chi2
    torch = _torch()
    mask = torch.as_tensor(b0_mask(cfp), device=meas.device)
    k = int(b0_mask(cfp).sum())
    if k < 2:
        raise ValueError("the noise bound needs at least two b=0 measurements")
    v = meas[:, mask].double().var(1, unbiased=True)
    return float(v.median()) / (chi2.median(k - 1) / (k - 1))


def bingham_learned_lm(meas, cfp, s0, net, lm_iters=200, fallback=True, factor=1.5, stream=None):
    """Learned start + psi scan + LM, with the GridSearch fallback.
    Returns (params, cost, flagged) on the device; flagged marks the voxels
    that went through the fallback.  lm_iters <= 0 returns the learned start
    (cost None)."""
    torch = _torch()
    n, nmeas = meas.shape
    d_cfp = _dev(cfp)
    m = capi.Model("NODDI_Bingham")
    spec = m.spec()
    p = _dev(np.tile(np.asarray(spec["p_init"], dtype=np.float32), (n, 1)))
    net.init_params(meas, s0[:, 0] if s0.dim() == 2 else s0, p, stream=stream)
    psi = "search[6]=(%s)\n" % ",".join("%.9g" % (np.pi * k / PSI_SCAN) for k in range(PSI_SCAN))
    scan = _model("NODDI_Bingham", grid=psi)
    scan.gridsearch(meas, p, d_cfp, s0, stream=stream)
    cost = torch.empty(n, device=meas.device)
    flagged = torch.zeros(n, dtype=torch.bool, device=meas.device)
    if lm_iters <= 0:  # the learned start alone
        return p, None, flagged
    m.levmar(meas, p, d_cfp, s0, max_iters=lm_iters, cost=cost, stream=stream)
    if fallback:
        bound = factor * (nmeas - 7) * noise_sigma2(meas, cfp)
        flagged = cost > bound
        idx = torch.nonzero(flagged).flatten()
        if idx.numel():
            pg, cg = bingham_gridsearch_lm(meas[idx].contiguous(), cfp, s0[idx].contiguous(), lm_iters=lm_iters,
                                           stream=stream)
            better = cg < cost[idx]
            sel = idx[better]
            p[sel] = pg[better]
            cost[sel] = cg[better]
    return p, cost, flagged
