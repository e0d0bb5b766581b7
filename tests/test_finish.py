"""Derived maps of the <Model>_finish.sh steps (cdm_sample_mean, cdm_dyads,
cdm_noddi_maps).  CPU tests pin the oracle (oracle/finish_ref.py) with
closed forms and invariants; GPU tests compare the kernels with it."""
import numpy as np
import pytest

from oracle import finish_ref as fr


def _stick_samples(n, S, seed):
    rng = np.random.default_rng(seed)
    x = np.empty((n, 7, S))
    x[:, 0] = rng.uniform(0, 0.5, (n, S))               # fiso
    x[:, 1] = rng.uniform(0.2, 0.9, (n, S))             # fintra
    x[:, 2] = rng.uniform(1, 40, (n, 1)) * rng.uniform(0.9, 1.1, (n, S))  # kappa
    x[:, 3] = x[:, 2] * rng.uniform(0, 0.9, (n, S))     # beta <= kappa
    th0, ph0 = rng.uniform(0, np.pi, (n, 1)), rng.uniform(-np.pi, np.pi, (n, 1))
    x[:, 4] = th0 + 0.2 * rng.normal(size=(n, S))
    x[:, 5] = ph0 + 0.2 * rng.normal(size=(n, S))
    x[:, 6] = rng.uniform(0, np.pi, (n, 1)) + 0.2 * rng.normal(size=(n, S))  # psi
    return x.astype(np.float32)


def test_oracle_dyads_of_identical_directions():
    th, ph = 0.7, -2.1
    s = np.zeros((1, 2, 9))
    s[0, 0], s[0, 1] = th, ph
    d, disp = fr.dyads(s, 0, 1)
    n = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    np.testing.assert_allclose(d[0], n * np.sign(n[2]), atol=1e-12)
    assert abs(disp[0]) < 1e-12


@pytest.mark.parametrize("a", [0.3, 0.8])
def test_oracle_dyads_of_two_orthogonal_directions(a):
    # a fraction a along z (theta 0), 1 - a along x (theta pi/2, phi 0):
    # eigenvalues a and 1 - a
    k = 10
    m = int(round(a * k))
    s = np.zeros((1, 2, k))
    s[0, 0, m:] = np.pi / 2
    d, disp = fr.dyads(s, 0, 1)
    want = np.array([0, 0, 1.0]) if a > 0.5 else np.array([1.0, 0, 0])
    np.testing.assert_allclose(np.abs(d[0]), want, atol=1e-12)
    np.testing.assert_allclose(disp[0], 1 - max(m / k, 1 - m / k), atol=1e-12)


def test_oracle_fanning_direction_invariants():
    rng = np.random.default_rng(4)
    th, ph, psi = rng.uniform(0, np.pi, 50), rng.uniform(-np.pi, np.pi, 50), rng.uniform(0, np.pi, 50)
    f = fr.directions(th, ph, psi)
    n = fr.directions(th, ph)
    np.testing.assert_allclose(np.linalg.norm(f, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.sum(f * n, axis=1), 0.0, atol=1e-12)  # fanning is across the fibre
    np.testing.assert_allclose(fr.directions(th, ph, 0 * psi), np.stack([-np.sin(ph), np.cos(ph), 0 * ph], 1),
                               atol=1e-12)
    dn_dth = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], 1)
    np.testing.assert_allclose(fr.directions(th, ph, 0 * psi + np.pi / 2), -dn_dth, atol=1e-12)


def test_oracle_noddi_closed_forms():
    s = np.zeros((3, 2, 4))
    s[:, 0] = np.array([1.0, 2.0, 5.0])[:, None]  # kappa
    s[:, 1] = np.array([0.5, 1.0, 0.0])[:, None]  # beta
    od = fr.noddi_maps(s, 0)
    np.testing.assert_allclose(od[0, 0], 0.5, atol=1e-12)  # atan(1) * 2/pi
    m = fr.noddi_maps(s, 0, 1)
    np.testing.assert_allclose(m[:, 0], [0.5, 1.0, 5.0])  # k2 = kappa - beta
    np.testing.assert_allclose(m[1, 4], 0.5, atol=1e-12)  # beta == k2 -> DA = 1/2
    np.testing.assert_allclose(m[2, 4], 0.0, atol=1e-12)  # beta = 0 -> DA = 0
    np.testing.assert_allclose(m[2, 3], m[2, 2], atol=1e-12)  # k2 == kappa -> ODItot == ODIs
    np.testing.assert_allclose(m[:, 1], np.arctan(1 / m[:, 0]) * 2 / np.pi, rtol=1e-6)


@pytest.mark.gpu
def test_sample_mean_matches_oracle():
    import torch
    from cudimot_amd import capi
    from tests.helpers import dev, host
    x = _stick_samples(3001, 50, 1)
    out = torch.empty((3001, 7), device="cuda")
    capi.sample_mean(dev(x), out)
    capi.sync()
    np.testing.assert_allclose(host(out), fr.sample_mean(x), rtol=1e-6, atol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("psi", [-1, 6])
def test_dyads_match_oracle(psi):
    import torch
    from cudimot_amd import capi
    from tests.helpers import dev, host
    n = 2049
    x = _stick_samples(n, 50, 2)
    d, disp = torch.empty((n, 3), device="cuda"), torch.empty(n, device="cuda")
    capi.dyads(dev(x), 4, 5, d, disp, psi=psi)
    capi.sync()
    wd, wdisp = fr.dyads(x, 4, 5, None if psi < 0 else psi)
    g = host(d)
    np.testing.assert_allclose(np.abs(np.sum(g * wd, axis=1)), 1.0, atol=2e-5)
    clear = np.abs(wd[:, 2]) > 1e-3
    np.testing.assert_array_equal(np.sign(g[clear, 2]), np.sign(wd[clear, 2]))
    np.testing.assert_allclose(host(disp), wdisp, atol=2e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("beta", [-1, 3])
def test_noddi_maps_match_oracle(beta):
    import torch
    from cudimot_amd import capi
    from tests.helpers import dev, host
    n = 1500
    x = _stick_samples(n, 50, 3)
    out = torch.empty((n, 1 if beta < 0 else 5), device="cuda")
    capi.noddi_maps(dev(x), 2, out, beta=beta)
    capi.sync()
    np.testing.assert_allclose(host(out), fr.noddi_maps(x, 2, None if beta < 0 else beta), rtol=2e-6, atol=1e-7)


def test_finish_entry_points_reject_bad_indices():
    from cudimot_amd import capi

    class F:
        shape = (10, 7, 5)

        def data_ptr(self):
            return 8
    with pytest.raises(capi.CdmError, match="out of range"):
        capi.dyads(F(), 7, 5, F())
    with pytest.raises(capi.CdmError, match="out of range"):
        capi.noddi_maps(F(), 2, F(), beta=9)
