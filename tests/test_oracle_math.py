"""Pins of the oracle's forward models by mathematics alone (no compiled
reference): closed forms at special points, invariances, reductions of one
model to another, and direct quadrature over the sphere.  Each check is
chosen so that a plausible mistake in a forward model fails it: a wrong
angle convention, a dropped factor or term, swapped parameters, a wrong
compartment weight or diffusivity.

What each check fixes:
  * Ball_k_Sticks: the ball alone at f = 0; each stick's term at a gradient
    placed at a known angle to its axis, built from the standard spherical
    unit vector of (theta, phi); linearity in the fractions; invariance
    under permuting the sticks, under a joint rotation of the gradients and
    the sticks, and under the antipodal map (pi - theta, phi + pi);
  * DiffusionTensor: S0 exp(-b g.(D g)) with D assembled from its six
    parameters (Dxx, Dxy, Dxz, Dyy, Dyz, Dzz), exact recovery of D by
    log-linear least squares, and eigen-direction signals;
  * NODDI_Watson / NODDI_Bingham: the isotropic compartment exp(-b 3e-3);
    linearity in fiso; fintra = 0 makes the hindered tensor isotropic
    (exp(-b 1.7e-3)); fintra = 1 leaves the intra-axonal signal (pinned by
    quadrature and eigenvalues in test_oracle.py); at intermediate fintra
    the hindered signal equals exp(-b g^T D_h g) with D_h the
    orientation-averaged tensor of sticks with tortuous perpendicular
    diffusivity Dpar (1 - fintra), by quadrature of the Watson density
    (Watson) and, for Bingham, against the finite differences of an
    independent saddle-point evaluation (bisection in extended precision)
    and by quadrature where the saddle-point approximation is accurate;
  * NODDI_Bingham_2Fans: fFan2 = 0 (1) is NODDI_Bingham on fan 1 (2);
    swapping the fans with fFan2 -> 1 - fFan2 leaves the signal; linearity
    in fFan2;
  * the NODDI analytic derivatives (fiso, and fintra without the
    tortuosity term, as the reference) from the compartments, and the
    central-difference ones against a Richardson-extrapolated derivative.
"""
import numpy as np
import pytest

from oracle import bind
from tests import helpers as H

DISO, DPAR = 3.0e-3, 1.7e-3


def unit(th, ph):
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)


def perp(n, rng):
    """A unit vector normal to n."""
    u = np.cross(n, rng.normal(size=3))
    return u / np.linalg.norm(u)


def angles(v):
    v = v / np.linalg.norm(v)
    return np.arccos(np.clip(v[2], -1, 1)), np.arctan2(v[1], v[0])


def rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


# ------------------------------------------------------------ Ball_k_Sticks

def _stick_cfp(n_axis, rng, bvals=(700.0, 1000.0, 2000.0, 3000.0)):
    """Gradients at known angles to n_axis: along it, normal to it, and at
    angles alpha in between; returns (cfp, cos^2 alpha per row)."""
    u = perp(n_axis, rng)
    alphas = [0.0, np.pi / 2, 0.3, 0.9, 1.3, 2.2, 2.9]
    rows, c2 = [], []
    for b in bvals:
        for a in alphas:
            g = np.cos(a) * n_axis + np.sin(a) * u
            rows.append(np.append(g, b))
            c2.append(np.cos(a) ** 2)
    rows.append([0.0, 0.0, 0.0, 0.0])
    c2.append(0.0)
    return np.array(rows), np.array(c2)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_ball_sticks_closed_forms(k):
    model = {1: "Ball_1_Stick", 2: "Ball_2_Sticks", 3: "Ball_3_Sticks"}[k]
    o = bind.OracleModel(model)
    rng = np.random.default_rng(100 + k)
    for _ in range(12):
        s0, d = rng.uniform(500, 1500), rng.uniform(0.3e-3, 3e-3)
        th = rng.uniform(0.1, np.pi - 0.1, k)
        ph = rng.uniform(-np.pi, np.pi, k)
        f = rng.dirichlet(np.ones(k + 1))[:k]
        p = np.zeros(2 + 3 * k)
        p[0], p[1] = s0, d
        for j in range(k):
            p[2 + 3 * j:5 + 3 * j] = (f[j], th[j], ph[j])
        for j in range(k):
            cfp, c2 = _stick_cfp(unit(th[j], ph[j]), rng)
            b = cfp[:, 3]
            # only stick j: its term at known angles, the ball for the rest
            q = p.copy()
            for i in range(k):
                if i != j:
                    q[2 + 3 * i] = 0.0
            want = s0 * ((1 - f[j]) * np.exp(-b * d) + f[j] * np.exp(-b * d * c2))
            np.testing.assert_allclose(o.signal(q, cfp)[0], want, rtol=1e-13)
        # the ball alone
        cfp = H.ACQ[H.MODEL_ACQ[model]]()
        q = p.copy()
        q[2::3] = 0.0
        np.testing.assert_allclose(o.signal(q, cfp)[0], s0 * np.exp(-cfp[:, 3] * d), rtol=1e-13)
        # linearity in the fractions: S(f) = S(0) + sum_j (S(f_j e_j) - S(0))
        base = o.signal(q, cfp)[0]
        acc = base.copy()
        for j in range(k):
            qj = q.copy()
            qj[2 + 3 * j] = f[j]
            acc += o.signal(qj, cfp)[0] - base
        np.testing.assert_allclose(o.signal(p, cfp)[0], acc, rtol=1e-12)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_ball_sticks_invariances(k):
    model = {1: "Ball_1_Stick", 2: "Ball_2_Sticks", 3: "Ball_3_Sticks"}[k]
    o = bind.OracleModel(model)
    rng = np.random.default_rng(200 + k)
    cfp = H.ACQ[H.MODEL_ACQ[model]]()
    for _ in range(10):
        p = H.truth(model, 1, int(rng.integers(1 << 30)))[0]
        s = o.signal(p, cfp)[0]
        # antipodal axes
        q = p.copy()
        for j in range(k):
            q[3 + 3 * j] = np.pi - p[3 + 3 * j]
            q[4 + 3 * j] = p[4 + 3 * j] + np.pi
        np.testing.assert_allclose(o.signal(q, cfp)[0], s, rtol=1e-12)
        # joint rotation of gradients and sticks
        R = rot(rng)
        cr = cfp.copy()
        cr[:, :3] = cfp[:, :3] @ R.T
        q = p.copy()
        for j in range(k):
            q[3 + 3 * j], q[4 + 3 * j] = angles(R @ unit(p[3 + 3 * j], p[4 + 3 * j]))
        np.testing.assert_allclose(o.signal(q, cr)[0], s, rtol=1e-11)
        # the sticks are interchangeable in the signal
        if k > 1:
            perm = rng.permutation(k)
            q = p.copy()
            for j, pj in enumerate(perm):
                q[2 + 3 * j:5 + 3 * j] = p[2 + 3 * pj:5 + 3 * pj]
            np.testing.assert_allclose(o.signal(q, cfp)[0], s, rtol=1e-12)


# -------------------------------------------------------------- DiffusionTensor

def _tensor(p):
    return np.array([[p[1], p[2], p[3]], [p[2], p[4], p[5]], [p[3], p[5], p[6]]])


def test_dti_closed_form_and_log_linear_recovery():
    o = bind.OracleModel("DiffusionTensor")
    rng = np.random.default_rng(7)
    cfp = H.ACQ["M64"]()
    g, b = cfp[:, :3], cfp[:, 3]
    for _ in range(20):
        p = H.truth("DiffusionTensor", 1, int(rng.integers(1 << 30)))[0]
        p[0] = rng.uniform(500, 1500)
        D = _tensor(p)
        s = o.signal(p, cfp)[0]
        want = p[0] * np.exp(-b * np.einsum("mi,ij,mj->m", g, D, g))
        np.testing.assert_allclose(s, want, rtol=1e-13)
        # log-linear least squares over the 6 tensor entries and log S0
        dw = b > 0
        A = np.concatenate([np.ones((dw.sum(), 1)),
                            -b[dw, None] * np.stack([g[dw, 0] ** 2, 2 * g[dw, 0] * g[dw, 1], 2 * g[dw, 0] * g[dw, 2],
                                                      g[dw, 1] ** 2, 2 * g[dw, 1] * g[dw, 2], g[dw, 2] ** 2], 1)], 1)
        x = np.linalg.lstsq(A, np.log(s[dw]), rcond=None)[0]
        np.testing.assert_allclose(np.exp(x[0]), p[0], rtol=1e-9)
        np.testing.assert_allclose(x[1:], p[1:], rtol=1e-7, atol=1e-12)
        # along an eigenvector the signal is S0 exp(-b lambda)
        lam, V = np.linalg.eigh(D)
        ce = np.array([np.append(V[:, i], 1500.0) for i in range(3)])
        np.testing.assert_allclose(o.signal(p, ce)[0], p[0] * np.exp(-1500.0 * lam), rtol=1e-12)


# ------------------------------------------------------------------- NODDI

def _sphere(mu, nt=200, na=400):
    a = np.array([1.0, 0, 0]) if abs(mu[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(mu, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(mu, e1)
    t, w = np.polynomial.legendre.leggauss(nt)
    az = np.linspace(0, 2 * np.pi, na, endpoint=False)
    T, A = np.meshgrid(t, az, indexing="ij")
    W = np.repeat(w[:, None], na, 1)
    s = np.sqrt(1 - T ** 2)
    N = T[..., None] * mu + s[..., None] * (np.cos(A)[..., None] * e1 + np.sin(A)[..., None] * e2)
    return N.reshape(-1, 3), W.ravel()


def _hindered_tensor(N, dens, dperp):
    """Orientation average of the stick tensors dperp I + (Dpar - dperp) n n^T."""
    w = dens / dens.sum()
    M2 = np.einsum("k,ki,kj->ij", w, N, N)
    return dperp * np.eye(3) + (DPAR - dperp) * M2


def _noddi_setup(model, rng):
    th, ph, psi = rng.uniform(0.2, 2.9), rng.uniform(-np.pi, np.pi), rng.uniform(0, np.pi)
    mu = unit(th, ph)
    st, ct, sp, cp, ss, cs = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph), np.sin(psi), np.cos(psi)
    fan = np.array([-cs * sp - ct * cp * ss, cp * cs - ct * sp * ss, st * ss])
    return th, ph, psi, mu, fan


def _params(model, fiso, fin, kappa, beta, th, ph, psi):
    if model == "NODDI_Watson":
        return np.array([fiso, fin, kappa, th, ph])
    return np.array([fiso, fin, kappa, beta, th, ph, psi])


@pytest.mark.parametrize("model", ["NODDI_Watson", "NODDI_Bingham"])
def test_noddi_compartment_limits(model):
    o = bind.OracleModel(model)
    rng = np.random.default_rng(31)
    cfp = H.ACQ["M99"]()
    b = cfp[:, 3]
    s0 = np.array([[1234.5]])
    for _ in range(8):
        th, ph, psi, _, _ = _noddi_setup(model, rng)
        kappa = rng.uniform(0.5, 60)
        beta = rng.uniform(0, 0.9) * min(kappa, 32)
        fin, fiso = rng.uniform(0.1, 0.9), rng.uniform(0.05, 0.5)
        # isotropic compartment alone
        s = o.signal(_params(model, 1.0, fin, kappa, beta, th, ph, psi), cfp, s0)[0]
        np.testing.assert_allclose(s, 1234.5 * np.exp(-b * DISO), rtol=1e-6)
        # fintra = 0: tortuosity gives dperp = Dpar, an isotropic hindered tensor
        s = o.signal(_params(model, 0.0, 0.0, kappa, beta, th, ph, psi), cfp, s0)[0]
        np.testing.assert_allclose(s, 1234.5 * np.exp(-b * DPAR), rtol=1e-6)
        # linear in fiso
        s_a = o.signal(_params(model, 0.0, fin, kappa, beta, th, ph, psi), cfp, s0)[0]
        s_b = o.signal(_params(model, 1.0, fin, kappa, beta, th, ph, psi), cfp, s0)[0]
        s = o.signal(_params(model, fiso, fin, kappa, beta, th, ph, psi), cfp, s0)[0]
        np.testing.assert_allclose(s, (1 - fiso) * s_a + fiso * s_b, rtol=1e-12)


def _decompose(o, model, kappa, beta, th, ph, psi, fin, cfp):
    """Hindered signal E at fintra = fin from the signals at fin and at 1
    (S/S0 = (1 - fin) E + fin I with fiso = 0)."""
    one = np.array([[1.0]])
    s = o.signal(_params(model, 0.0, fin, kappa, beta, th, ph, psi), cfp, one)[0]
    i = o.signal(_params(model, 0.0, 1.0, kappa, beta, th, ph, psi), cfp, one)[0]
    return (s - fin * i) / (1 - fin), i


def test_watson_hindered_vs_quadrature():
    o = bind.OracleModel("NODDI_Watson")
    rng = np.random.default_rng(41)
    cfp = H.ACQ["M99"]()
    for kappa in (0.3, 1.0, 2.5, 6.0, 16.0, 40.0):
        th, ph, psi, mu, _ = _noddi_setup("NODDI_Watson", rng)
        N, W = _sphere(mu)
        dens = W * np.exp(kappa * (N @ mu) ** 2)
        for fin in (0.2, 0.55, 0.85):
            E, _ = _decompose(o, "NODDI_Watson", kappa, 0.0, th, ph, psi, fin, cfp)
            Dh = _hindered_tensor(N, dens, DPAR * (1 - fin))
            want = np.exp(-cfp[:, 3] * np.einsum("mi,ij,mj->m", cfp[:, :3], Dh, cfp[:, :3]))
            # the reference's Dawson approximation is good to ~2e-7
            np.testing.assert_allclose(E, want, rtol=2e-6, err_msg="kappa %g fintra %g" % (kappa, fin))


def _hyp_saddle_ld(l):
    """Second-order saddle-point 1F1(1/2; 3/2; -diag(l)) (Kume & Wood), the
    saddle point by bisection in extended precision."""
    l = np.asarray(l, dtype=np.longdouble)
    if np.all(l == 0):
        return np.longdouble(1)
    lo, hi = l.min() - np.longdouble(2), l.min() - np.longdouble(0.25)
    for _ in range(200):
        mid = (lo + hi) / 2
        if 0.5 * np.sum(1 / (l - mid)) > 1:
            hi = mid
        else:
            lo = mid
    t = (lo + hi) / 2
    y = 1 / (l - t)
    K2, K3, K4 = 0.5 * np.sum(y ** 2), np.sum(y ** 3), 3 * np.sum(y ** 4)
    T2 = K4 / (8 * K2 ** 2) - 5 * K3 ** 2 / (24 * K2 ** 3)
    return 0.25 * np.sqrt(2 * np.prod(y) / K2) * np.exp(T2 - t)


def _bingham_moments_saddle(kappa, beta, delta=5e-5):
    """E[(n.mu)^2], E[(n.fan)^2] of the Bingham density as the derivatives of
    log 1F1 in kappa and beta, by central differences (step delta) of the
    independent saddle-point evaluation."""
    F = lambda k, b: float(_hyp_saddle_ld(-np.array([0.0, b, k])))
    f0 = F(kappa, beta)
    m_mu = (F(kappa + delta, beta) - F(kappa - delta, beta)) / (2 * delta) / f0
    m_fan = (F(kappa, beta + delta) - F(kappa, beta - delta)) / (2 * delta) / f0
    return m_mu, m_fan


def test_bingham_hindered_vs_saddle_moments_and_quadrature():
    """Hindered tensor D_h = dperp I + (Dpar - dperp) E[n n^T]; the Bingham
    second moment has eigenvalues E[(n.mu)^2] along the fibre, E[(n.fan)^2]
    along the fanning direction and the rest along their cross product.
    Against the saddle-point moments (as the reference defines them) to
    1e-6, and against quadrature of the Bingham density where the
    saddle-point approximation is accurate (large kappa)."""
    o = bind.OracleModel("NODDI_Bingham")
    rng = np.random.default_rng(43)
    cfp = H.ACQ["M99"]()
    g, b = cfp[:, :3], cfp[:, 3]
    for kappa, beta, qtol in ((2.0, 1.0, None), (6.0, 5.0, None), (12.0, 3.0, None), (25.0, 8.0, 2e-3),
                              (40.0, 30.0, 2e-3), (60.0, 1.0, 5e-4)):
        th, ph, psi, mu, fan = _noddi_setup("NODDI_Bingham", rng)
        x = np.cross(fan, mu)
        m_mu, m_fan = _bingham_moments_saddle(kappa, beta)
        M2 = m_mu * np.outer(mu, mu) + m_fan * np.outer(fan, fan) + (1 - m_mu - m_fan) * np.outer(x, x)
        N, W = _sphere(mu, 300, 600)
        dens = W * np.exp(kappa * (N @ mu) ** 2 + beta * (N @ fan) ** 2)
        for fin in (0.3, 0.7):
            E, _ = _decompose(o, "NODDI_Bingham", kappa, beta, th, ph, psi, fin, cfp)
            dperp = DPAR * (1 - fin)
            Dh = dperp * np.eye(3) + (DPAR - dperp) * M2
            want = np.exp(-b * np.einsum("mi,ij,mj->m", g, Dh, g))
            np.testing.assert_allclose(E, want, rtol=1e-6, err_msg="kappa %g beta %g" % (kappa, beta))
            if qtol:
                Dq = _hindered_tensor(N, dens, dperp)
                wq = np.exp(-b * np.einsum("mi,ij,mj->m", g, Dq, g))
                np.testing.assert_allclose(E, wq, rtol=qtol, err_msg="quadrature kappa %g beta %g" % (kappa, beta))


def test_bingham_2fans_reductions():
    o2 = bind.OracleModel("NODDI_Bingham_2Fans")
    o1 = bind.OracleModel("NODDI_Bingham")
    rng = np.random.default_rng(47)
    cfp = H.ACQ["M99"]()
    s0 = np.array([[987.0]])
    for _ in range(10):
        t = H.truth("NODDI_Bingham_2Fans", 1, int(rng.integers(1 << 30)))[0]
        fiso = t[0]
        fan1, fan2 = t[2:8], t[8:14]
        single = lambda fan: o1.signal(np.concatenate([[fiso], fan]), cfp, s0)[0]
        two = lambda ff, a, b: o2.signal(np.concatenate([[fiso, ff], a, b]), cfp, s0)[0]
        np.testing.assert_allclose(two(0.0, fan1, fan2), single(fan1), rtol=1e-12)
        np.testing.assert_allclose(two(1.0, fan1, fan2), single(fan2), rtol=1e-12)
        ff = rng.uniform(0.05, 0.5)
        np.testing.assert_allclose(two(ff, fan1, fan2), two(1 - ff, fan2, fan1), rtol=1e-12)
        # linear in fFan2, with the isotropic part outside the mixture
        np.testing.assert_allclose(two(ff, fan1, fan2), (1 - ff) * two(0.0, fan1, fan2) + ff * two(1.0, fan1, fan2),
                                   rtol=1e-12)


@pytest.mark.parametrize("model", ["NODDI_Watson", "NODDI_Bingham"])
def test_noddi_derivatives(model):
    """Analytic entries: dS/dfiso = S0 (iso - aniso) exactly; dS/dfintra =
    S0 (1 - fiso) (I - E), the reference's form that leaves out the
    tortuosity dependence of E (it differs from the true derivative).  The
    central-difference entries (kappa, beta, angles) against a Richardson
    extrapolation of central differences of the signal."""
    o = bind.OracleModel(model)
    rng = np.random.default_rng(53)
    cfp = H.ACQ["M99"]()
    b = cfp[:, 3]
    s0 = np.array([[1000.0]])
    for _ in range(6):
        th, ph, psi, _, _ = _noddi_setup(model, rng)
        kappa = rng.uniform(2, 50)
        beta = rng.uniform(0.1, 0.8) * min(kappa, 32)
        fin, fiso = rng.uniform(0.2, 0.8), rng.uniform(0.05, 0.4)
        p = _params(model, fiso, fin, kappa, beta, th, ph, psi)
        d = o.derivs(p, cfp, s0)[0]
        E, I = _decompose(o, model, kappa, beta, th, ph, psi, fin, cfp)
        aniso = (1 - fin) * E + fin * I
        np.testing.assert_allclose(d[:, 0], 1000.0 * (np.exp(-b * DISO) - aniso), rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(d[:, 1], 1000.0 * (1 - fiso) * (I - E), rtol=1e-6, atol=1e-9)
        # the true fintra derivative has the tortuosity term: it is not the one used
        h = 1e-4
        full = (o.signal(p + h * np.eye(len(p))[1], cfp, s0)[0] - o.signal(p - h * np.eye(len(p))[1], cfp, s0)[0]) / (
            2 * h)
        assert np.abs(full - d[:, 1]).max() > 1e-3 * np.abs(full).max()
        for i in range(2, len(p)):
            e = np.eye(len(p))[i]

            def cd(hh):
                return (o.signal(p + hh * e, cfp, s0)[0] - o.signal(p - hh * e, cfp, s0)[0]) / (2 * hh)
            hh = 1e-3 * max(1.0, abs(p[i]))
            rich = (4 * cd(hh / 2) - cd(hh)) / 3
            scale = np.abs(rich).max() + 1e-12
            # Bingham's kappa and beta enter the hindered diffusivities through
            # central differences (step 5e-5) of the 1F1 normaliser, whose
            # ~1e-12 rounding leaves ~1e-11 relative noise in the signal; the
            # reference's step 1e-5 |p| then carries up to ~1e-3 of noise
            tol = 3e-3 if model == "NODDI_Bingham" and i in (2, 3) else 2e-4
            assert np.abs(d[:, i] - rich).max() <= tol * scale, (model, i, np.abs(d[:, i] - rich).max() / scale)
