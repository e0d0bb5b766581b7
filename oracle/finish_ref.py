"""ORACLE (test infrastructure only; never imported by the product).

CPU restatement, in numpy FP64, of the derived maps the reference's
<Model>_finish.sh scripts compute from the merged MCMC samples, i.e. of
cudimot_amd's cdm_sample_mean / cdm_dyads / cdm_noddi_maps:

- sample means: fslmaths -Tmean of each Param_k_samples volume
  (e.g. mymodels/NODDI_Watson/NODDI_Watson_finish.sh:30-32);
- mean orientation: principal eigenvector of the mean dyadic tensor of the
  sampled directions and dispersion 1 - its eigenvalue (FSL
  make_dyadic_vectors, called at Ball_1_Stick_finish.sh:31; FSL is not in
  this image, so its published definition is restated), and the Bingham
  fanning direction of utils/getFanningOrientation.cc:39-76 (direction
  formula :49-51, mean tensor :54, largest eigenvalue :55-64);
- NODDI indices: OD = 2/pi atan(1/mean kappa) (NODDI_Watson_finish.sh:35);
  k2 = kappa - beta per sample (NODDI_Bingham_finish.sh:32), ODIp, ODIs,
  ODItot and DA (:46-56).

Eigenvector signs are arbitrary in the reference (they come from NEWMAT's
solver): the oracle and the kernels report the sign with z >= 0.
samples: [nvox][nparams][nsamples].
"""
import numpy as np

TWO_DIV_PI = 0.636619772367581


def sample_mean(samples):
    return samples.astype(np.float64).mean(axis=2)


def directions(th, ph, psi=None):
    """Unit vectors per sample: the fibre direction, or the fanning
    direction when psi is given (getFanningOrientation.cc:49-51)."""
    th, ph = np.asarray(th, np.float64), np.asarray(ph, np.float64)
    if psi is None:
        return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    psi = np.asarray(psi, np.float64)
    return np.stack([-np.cos(psi) * np.sin(ph) - np.cos(th) * np.cos(ph) * np.sin(psi),
                     np.cos(ph) * np.cos(psi) - np.cos(th) * np.sin(ph) * np.sin(psi),
                     np.sin(th) * np.sin(psi)], -1)


def canonical_sign(v):
    key = np.where(v[..., 2] != 0, v[..., 2], np.where(v[..., 1] != 0, v[..., 1], v[..., 0]))
    return v * np.where(key < 0, -1.0, 1.0)[..., None]


def dyads(samples, th, ph, psi=None):
    """(dyads [nvox][3], dispersion [nvox])."""
    d = directions(samples[:, th, :], samples[:, ph, :], None if psi is None else samples[:, psi, :])
    t = np.einsum("vsi,vsj->vij", d, d) / d.shape[1]
    w, v = np.linalg.eigh(t)
    return canonical_sign(v[:, :, -1]), 1.0 - np.abs(w).max(axis=1)


def noddi_maps(samples, kappa, beta=None):
    """Watson: [nvox][1] OD.  Bingham: [nvox][5] mean k2, ODIp, ODIs,
    ODItot, DA."""
    s = samples.astype(np.float64)
    mk = s[:, kappa, :].mean(axis=1)
    if beta is None:
        return (np.arctan(1.0 / mk) * TWO_DIV_PI)[:, None]
    mb = s[:, beta, :].mean(axis=1)
    mk2 = (s[:, kappa, :] - s[:, beta, :]).mean(axis=1)
    return np.stack([mk2, np.arctan(1.0 / mk2) * TWO_DIV_PI, np.arctan(1.0 / mk) * TWO_DIV_PI,
                     np.arctan(np.abs(np.sqrt((1.0 / mk) * (1.0 / mk2)))) * TWO_DIV_PI,
                     np.arctan(mb / mk2) * TWO_DIV_PI], 1)
