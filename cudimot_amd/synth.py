"""Seeded synthetic diffusion acquisitions, ground truths and Rician noise.

Used by the benchmark and the tests to build volumes of the shapes the
fitting path is measured on:
  * gradient directions uniform on the sphere (normalised Gaussian triplets)
    from the seed, b=0 volumes with direction (1, 0, 0) and b = 0;
  * Rician noise |S + sigma (n1 + i n2)| with sigma = S0 / SNR (SNR on b=0);
  * ground-truth parameter draws per workload.
The CFP table has four columns per measurement: gx, gy, gz, b (the
reference's bvecs (3) + bvals (1), Parameters.cu:117-120).
"""
import numpy as np


def acquisition(n_b0, shells, seed):
    """shells: list of (bvalue, ndirs).  Returns cfp [M, 4] float64."""
    rng = np.random.default_rng(seed)
    rows = [[1.0, 0.0, 0.0, 0.0]] * n_b0
    for b, nd in shells:
        g = rng.normal(size=(nd, 3))
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        rows += [[x, y, z, float(b)] for x, y, z in g]
    return np.asarray(rows, dtype=np.float64)


def unit_sphere(rng, n):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def sph_angles(v):
    """(theta, phi) of unit vectors, theta in [0, pi], phi in (-pi, pi]."""
    th = np.arccos(np.clip(v[:, 2], -1.0, 1.0))
    ph = np.arctan2(v[:, 1], v[:, 0])
    return th, ph


def rician(signal, s0, snr, rng):
    sigma = (np.asarray(s0, dtype=np.float64) / snr).reshape(-1, 1)
    n1 = rng.normal(size=signal.shape) * sigma
    n2 = rng.normal(size=signal.shape) * sigma
    return np.sqrt((signal + n1) ** 2 + n2 ** 2)


# ------------------------------------------------------------ ground truths

def truth_ball_1_stick(n, seed):
    """S0 U[800,1200], d U[0.5,3]e-3, f U[0.1,0.8], orientation uniform."""
    rng = np.random.default_rng(seed)
    th, ph = sph_angles(unit_sphere(rng, n))
    return np.stack([rng.uniform(800, 1200, n), rng.uniform(0.5e-3, 3e-3, n), rng.uniform(0.1, 0.8, n), th, ph], 1)


def truth_ball_k_sticks(n, k, seed, s0=1000.0, min_angle_deg=30.0, d_range=(0.5e-3, 3e-3)):
    """Fractions (f1..fk, ball) ~ Dirichlet(1,...,1), sorted f1>=f2>=...;
    stick orientations uniform with pairwise angle >= min_angle."""
    rng = np.random.default_rng(seed)
    fr = rng.dirichlet(np.ones(k + 1), size=n)[:, :k]
    fr = -np.sort(-fr, axis=1)
    cosmax = np.cos(np.deg2rad(min_angle_deg))
    dirs = np.zeros((n, k, 3))
    for i in range(n):
        got = []
        while len(got) < k:
            v = unit_sphere(rng, 1)[0]
            if all(abs(v @ u) <= cosmax for u in got):
                got.append(v)
        dirs[i] = got
    cols = [np.full(n, s0), rng.uniform(*d_range, n)]
    for j in range(k):
        th, ph = sph_angles(dirs[:, j])
        cols += [fr[:, j], th, ph]
    return np.stack(cols, 1)


def truth_noddi_watson(n, seed):
    """fiso U[0,0.3], fintra U[0.1,0.9], ODI U[0.02,0.9] -> kappa =
    1/tan(ODI pi/2), orientation uniform."""
    rng = np.random.default_rng(seed)
    odi = rng.uniform(0.02, 0.9, n)
    th, ph = sph_angles(unit_sphere(rng, n))
    return np.stack([rng.uniform(0, 0.3, n), rng.uniform(0.1, 0.9, n), 1.0 / np.tan(odi * np.pi / 2), th, ph], 1)


def truth_noddi_bingham(n, seed):
    """fiso U[0,0.3], fintra U[0.1,0.9], kappa1 U[1,64], kappa2/kappa1 U[0,1]
    (kappa = kappa1, beta = kappa1 - kappa2), orientation uniform, psi U[0,pi)."""
    rng = np.random.default_rng(seed)
    k1 = rng.uniform(1, 64, n)
    k2 = k1 * rng.uniform(0, 1, n)
    th, ph = sph_angles(unit_sphere(rng, n))
    return np.stack([rng.uniform(0, 0.3, n), rng.uniform(0.1, 0.9, n), k1, k1 - k2, th, ph,
                     rng.uniform(0, np.pi, n)], 1)


def truth_dti(n, seed, s0=1000.0):
    """Random SPD tensors: eigenvalues U[0.2,3]e-3, random rotation."""
    rng = np.random.default_rng(seed)
    out = np.zeros((n, 7))
    for i in range(n):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        d = q @ np.diag(rng.uniform(0.2e-3, 3e-3, 3)) @ q.T
        out[i] = [s0, d[0, 0], d[0, 1], d[0, 2], d[1, 1], d[1, 2], d[2, 2]]
    return out


def truth_toymodel(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.5, 2.0, n), rng.uniform(0.2, 2.0, n)], 1)


def truth_noddi_bingham_2fans(n, seed):
    """Two Bingham fans: fiso U[0,0.3], fFan2 U[0.1,0.5], per fan the
    Bingham-NODDI draws above."""
    t1 = truth_noddi_bingham(n, seed)
    t2 = truth_noddi_bingham(n, seed + 1)
    ff = np.random.default_rng(seed + 2).uniform(0.1, 0.5, n)
    return np.concatenate([t1[:, :1], ff[:, None], t1[:, 1:], t2[:, 1:]], 1)


ACQUISITIONS = {
    # the measured workloads: (b=0 count, [(b, directions), ...], seed)
    "M64": (4, [(1000, 30), (2000, 30)], 11),
    "M99": (9, [(700, 30), (2000, 60)], 12),
    "M280": (10, [(1000, 90), (2000, 90), (3000, 90)], 13),
}

MODEL_ACQUISITION = {"Ball_1_Stick": "M64", "Ball_2_Sticks": "M280", "Ball_3_Sticks": "M280",
                     "DiffusionTensor": "M64", "NODDI_Watson": "M99", "NODDI_Watson_refnum": "M99",
                     "NODDI_Bingham": "M99",
                     "NODDI_Bingham_2Fans": "M99", "toymodel": None}


def model_acquisition(model):
    key = MODEL_ACQUISITION.get(model)
    if key is None:
        return np.linspace(0.0, 3.0, 20).reshape(-1, 1)
    n0, shells, seed = ACQUISITIONS[key]
    return acquisition(n0, shells, seed)


def clip_to_bounds(model, t):
    """Ground truths moved inside the shipped plug-ins' bounds and MCMC
    constraints (Bingham: beta <= 32 and beta < kappa - 0.1)."""
    t = t.copy()
    if model == "NODDI_Bingham":
        t[:, 2] = np.clip(t[:, 2], 0.2, 63.0)
        t[:, 3] = np.clip(np.minimum(t[:, 3], t[:, 2] - 0.2), 0.02, 31.0)
    elif model == "NODDI_Bingham_2Fans":
        for k, b in ((3, 4), (9, 10)):
            t[:, k] = np.clip(t[:, k], 0.2, 63.0)
            t[:, b] = np.clip(np.minimum(t[:, b], t[:, k] - 0.6), 0.02, 31.0)
    elif model in ("NODDI_Watson", "NODDI_Watson_refnum"):
        t[:, 2] = np.clip(t[:, 2], 0.01, 63.0)
    return t


def truth_for(model, n, seed):
    gen = {"Ball_1_Stick": truth_ball_1_stick,
           "Ball_2_Sticks": lambda n, s: truth_ball_k_sticks(n, 2, s),
           "Ball_3_Sticks": lambda n, s: truth_ball_k_sticks(n, 3, s),
           "DiffusionTensor": truth_dti, "NODDI_Watson": truth_noddi_watson,
           "NODDI_Watson_refnum": truth_noddi_watson,
           "NODDI_Bingham": truth_noddi_bingham, "NODDI_Bingham_2Fans": truth_noddi_bingham_2fans,
           "toymodel": truth_toymodel}[model]
    return clip_to_bounds(model, gen(n, seed))
