"""Mean-field coexistence theory (activemodelbplus.activedroplets), pinned by closed forms
(passive binodals, passive interface tension and tanh profile, Gibbs-Thomson), exact
symmetries, the pseudo-pressure identity of curved interfaces, and the digitised curves of
Tjhung, Nardini and Cates (2018) that the reference ships (tests/golden/tjhung2018): the
boundaries of reverse Ostwald ripening and the coexisting densities of droplets vs R."""
import os

import numpy as np
import pytest

from activemodelbplus import activedroplets as AD

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "tjhung2018")


def golden(name):
    x, y = np.genfromtxt(os.path.join(GOLDEN, name)).T
    o = np.argsort(x)
    return x[o], y[o]


@pytest.mark.parametrize("K,t,u", [(1.0, -0.25, 0.25), (2.0, -0.5, 0.3), (0.5, -1.0, 2.0)])
def test_passive_limits(K, t, u):
    """zeta = 2 lambda: the pseudodensity is phi, so the binodals are +-sqrt(-t/u) and the
    pseudo-tension is the equilibrium tension (2 sqrt 2 / 3) sqrt(K) |t|^{3/2} / u."""
    sigma = 2 * np.sqrt(2) / 3 * np.sqrt(K) * (-t) ** 1.5 / u
    for zeta, lamb in [(0.0, 0.0), (1.0, 0.5), (1e-10, 0.0)]:
        m = AD.ActiveModelBPlus(zeta, lamb, K, t, u)
        np.testing.assert_allclose(m.binodal_densities, [np.sqrt(-t / u), -np.sqrt(-t / u)], rtol=1e-9)
        if lamb == 0.0:
            assert m.surface_tension() == pytest.approx(sigma, rel=1e-8)


def test_passive_planar_profile_is_tanh():
    b = AD.ActiveModelBPlus(0, 0).binodal()
    x = np.linspace(-10, 10, 41)
    np.testing.assert_allclose(b(x), -np.tanh(x / np.sqrt(8.0)), atol=1e-4)


@pytest.mark.parametrize("zl", [-4.0, -2.0, -0.5, 1.0, 3.0])
def test_binodal_conditions_and_symmetry(zl):
    """Equal bulk chemical potential and pseudo-pressure; phi -> -phi maps zeta - 2 lambda to
    its negative and exchanges the phases."""
    m = AD.ActiveModelBPlus(zl, 0.0)
    a, b = m.binodal_densities
    assert m.free_energy_density(a, 1) == pytest.approx(m.free_energy_density(b, 1), abs=1e-12)
    assert m.bulk_pseudopressure(a) == pytest.approx(m.bulk_pseudopressure(b), abs=1e-12)
    a2, b2 = AD.ActiveModelBPlus(-zl, 0.0).binodal_densities
    assert a2 == pytest.approx(-b, abs=1e-10) and b2 == pytest.approx(-a, abs=1e-10)
    # the planar gradient vanishes at both binodals and is positive in between
    s = np.linspace(min(a, b), max(a, b), 101)
    q = m.planar_gradient_squared(s)
    assert abs(q[0]) < 1e-10 and abs(q[-1]) < 1e-10 and np.all(q[1:-1] > 0)
    phi0, phi1 = AD.binodal_densities(np.array([zl]))
    assert (phi0[0], phi1[0]) == pytest.approx((a, b), abs=1e-12)


def test_pseudotension_is_linear_in_lambda_and_symmetric():
    zl = -2.0
    s = [AD.ActiveModelBPlus(zl + 2 * l, l).surface_tension() for l in (-1.0, 0.0, 1.0)]
    assert s[1] - s[0] == pytest.approx(s[2] - s[1], rel=1e-9)
    # clusters of (zeta, lambda) are bubbles of (-zeta, -lambda)
    for zeta, lamb in [(-1, 0.5), (-4, -1), (2, 3)]:
        assert AD.ActiveModelBPlus(zeta, lamb).surface_tension("clusters") == pytest.approx(
            AD.ActiveModelBPlus(-zeta, -lamb).surface_tension("bubbles"), rel=1e-9)
    # Tjhung 2018: (-1, 0.5) has normal, (-4, -1) reversed Ostwald ripening of clusters
    assert AD.ActiveModelBPlus(-1, 0.5).surface_tension() > 0
    assert AD.ActiveModelBPlus(-4, -1).surface_tension() < 0


@pytest.mark.parametrize("which", ["clusters", "bubbles"])
def test_reversal_boundary_matches_tjhung2018(which):
    """Every digitised point (zeta, lambda) of the published boundary lies on the computed
    sigma = 0 line: lambda*(zeta - 2 lambda) within 0.006 (median 0.0013)."""
    z, l = np.genfromtxt(os.path.join(GOLDEN, f"tjhung2018_reversal_{which}.csv")).T
    err = np.array([AD.reversal_lambda(zi - 2 * li, which)[1] - li for zi, li in zip(z, l)])
    assert np.abs(err).max() < 0.006
    assert np.median(np.abs(err)) < 0.002
    # a mistake in the weight (dropping the lambda term's exp(alpha phi)) moves the line far off
    zl = -2.0
    m = AD.ActiveModelBPlus(zl, 0.0)
    hi, lo = np.sort(m.binodal_densities)[::-1]
    I0, I1 = m._tension_integrals(hi if which == "clusters" else lo)
    assert abs(AD.reversal_lambda(zl, which)[1] - (-zl * I0 / (2 * I1))) > 0.05


@pytest.mark.parametrize("zeta,lamb", [(-1, 0.5), (-4, -1)])
def test_droplet_densities_match_tjhung2018_theory(zeta, lamb):
    """phi(0) and phi(5R) of droplets with phi(R) = 0 against the published mean-field curves
    (every fourth digitised point, R >= 10).  The curves come from a sharp-interface
    treatment, so the diffuse profile departs from them as R approaches a few interface
    widths: phi0 of (-4, -1) by up to 0.03 at R = 13, within 0.01 from R = 20."""
    m = AD.ActiveModelBPlus(zeta, lamb, d=2)
    tol = {("phi0", -1): (0.002, 0.001), ("phi1", -1): (0.004, 0.001),
           ("phi0", -4): (0.035, 0.010), ("phi1", -4): (0.002, 0.002)}
    for which in ("phi0", "phi1"):
        R, p = golden(f"tjhung2018_zeta={zeta:g}_lamb={lamb:g}_d=2_{which}.csv")
        for lo, hi, k in [(10, 20, 0), (20, 41, 1)]:
            idx = np.nonzero((R >= lo) & (R < hi))[0][::4]
            err = [getattr(m.droplet(r), which) - v for r, v in zip(R[idx], p[idx])]
            assert np.abs(err).max() < tol[(which, zeta)][k], (which, lo, hi, err)


@pytest.mark.parametrize("d", [2, 3])
def test_droplet_gibbs_thomson(d):
    """Passive droplets: both densities shift by (d-1) sigma / (R dphi f''(phi_b)) to leading
    order in 1/R; the relative remainder halves when R doubles."""
    m = AD.ActiveModelBPlus(0, 0, d=d)
    sigma = m.surface_tension()
    rel = []
    for R in (50.0, 100.0):
        drop = m.droplet(R)
        shift = (d - 1) * sigma / (R * 2.0 * 0.5)
        rel.append(((drop.phi0 - 1) / shift - 1, (drop.phi1 + 1) / shift - 1))
        assert max(abs(r) for r in rel[-1]) < (d - 1) * 1.0 / R
    for k in range(2):
        assert 0.4 < rel[1][k] / rel[0][k] < 0.6


@pytest.mark.parametrize("zeta,lamb,d,R", [(-1, 0.5, 2, 15), (-4, -1, 2, 40), (-4, -1, 3, 15), (0, 1, 3, 40),
                                          (2, -1, 2, 15)])
def test_droplet_pseudopressure_identity(zeta, lamb, d, R):
    """P(phi0) - P(phi1) = (d-1) sigma_R / R holds exactly for the steady radial profile."""
    drop = AD.ActiveModelBPlus(zeta, lamb, d=d).droplet(R)
    assert drop(R) == pytest.approx(0.0, abs=1e-9)
    assert abs(drop(0.0, 1)) < 1e-9 and abs(drop(drop.L, 1)) < 1e-9
    assert drop.pseudopressure_drop == pytest.approx(drop.surface_tension_pseudopressure_drop, rel=1e-7)


def test_droplet_identity_tells_lambda_from_zeta():
    """(-1, 0.5) and (-4, -1) share zeta - 2 lambda (same planar binodals): a droplet solved
    for one fails the other's pressure identity, so the identity checks the curvature terms."""
    a = AD.ActiveModelBPlus(-1, 0.5, d=2).droplet(15)
    b = AD.ActiveModelBPlus(-4, -1, d=2).droplet(15)
    a.field_theory, b.field_theory = b.field_theory, a.field_theory
    for drop in (a, b):
        assert abs(drop.pseudopressure_drop - drop.surface_tension_pseudopressure_drop) > 0.01


def test_droplet_continuation_and_large_radius():
    m = AD.ActiveModelBPlus(-4, -1, d=2)
    d40 = m.droplet(40)
    d35 = m.droplet(35, guess=d40)
    assert d35.phi0 == pytest.approx(m.droplet(35).phi0, abs=1e-7)
    # densities approach the planar binodals as 1/R
    ex = [np.array([d.phi0, d.phi1]) - m.binodal_densities for d in (m.droplet(200), m.droplet(400))]
    np.testing.assert_array_less(np.abs(ex[1]), 0.01)
    assert np.all((0.4 < ex[1] / ex[0]) & (ex[1] / ex[0] < 0.6))


def test_argument_errors():
    with pytest.raises(ValueError):
        AD.ActiveModelBPlus(0, 0, t=0.25)
    with pytest.raises(ValueError):
        AD.ActiveModelBPlus(0, 0).droplet(10, domain_size=5)
    with pytest.raises(ValueError):
        AD.ActiveModelBPlus(0, 0).surface_tension("drops")
    with pytest.raises(ValueError):
        AD.ActiveModelBPlus(0, 0).pseudodensity(0.5, derivative=2)
