"""Fluctuation-dissipation on the lattice: in the Gaussian model (b = c = lamb = zeta = 0,
a > 0) the noisy explicit-Euler scheme has the exact stationary spectrum

    <|phi_k|^2> / N = T / (A(k) (1 - dt Gamma(k) A(k) / 2)),
    A = a - kappa L3(k),   Gamma = D8(kx)^2 + D8(ky)^2            (unit spacing)

with D8(k) = 2 [4/5 sin k - 1/5 sin 2k + 4/105 sin 3k - 1/280 sin 4k] and
L3(k) = -6 + 4 (cos kx + cos ky) - 2 cos kx cos ky, because the noise and the passive
current share D8 (reference: src/integrator.cu:203,259-260).  The modes with Gamma = 0 get
neither dynamics nor noise and stay at zero.  Measured from time averages of the
integrator's field (Philox noise on the device), in both precisions; the dt correction
reaches factors 1.6 and 2.1 at large k, so the continuum spectrum T / A fails the top shell."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 64


def d8(k):
    return 2 * (4 / 5 * np.sin(k) - 1 / 5 * np.sin(2 * k) + 4 / 105 * np.sin(3 * k) - 1 / 280 * np.sin(4 * k))


def theory(a, kappa, T, dt):
    k = 2 * np.pi * np.fft.fftfreq(N)
    kx, ky = k[None, :], k[:, None]
    gamma = d8(kx) ** 2 + d8(ky) ** 2
    l3 = -6 + 4 * (np.cos(kx) + np.cos(ky)) - 2 * np.cos(kx) * np.cos(ky)
    A = a - kappa * l3
    with np.errstate(divide="ignore", invalid="ignore"):
        s = T / (A * (1 - dt * gamma * A / 2))
    return s, gamma, A, np.hypot(kx, ky)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("a,kappa,T,dt", [(1.0, 1.0, 0.1, 0.01), (0.5, 0.5, 0.05, 0.03)])
def test_stationary_spectrum(precision, a, kappa, T, dt):
    import activemodelbplus.integrator as amb

    s_th, gamma, A, kmag = theory(a, kappa, T, dt)
    live = gamma > 1e-12
    assert np.max(dt * gamma * A) < 2  # stable
    sim = amb.Integrator(np.zeros((N, N)), amb.Stencil(dt, 1, 1), amb.Model(a, 0, 0, kappa, 0, 0, T), seed=11,
                         precision=precision)
    sim.run(int(round(600 / dt)))  # equilibrate: the slowest live mode relaxes in 100-200 time units
    acc = np.zeros((N, N))
    samples, gap = 400, int(round(5 / dt))
    for _ in range(samples):
        sim.run(gap)
        acc += np.abs(np.fft.fft2(sim.field)) ** 2
    s = acc / samples / N ** 2
    ratio = s[live] / s_th[live]
    print(f"precision {precision}, a={a}, kappa={kappa}, T={T}, dt={dt}: mean ratio {ratio.mean():.4f}, "
          f"dt factor up to {np.max(1 / (1 - dt * gamma[live] * A[live] / 2)):.2f}", flush=True)
    assert abs(ratio.mean() - 1) < 0.01
    # shells in |k|: each within 4% (10% for the lowest shell, whose modes are the slowest
    # and so the least sampled)
    edges = np.linspace(0, kmag[live].max() + 1e-9, 9)
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = live & (kmag >= lo) & (kmag < hi)
        if sel.sum() >= 16:
            r = s[sel].mean() / s_th[sel].mean()
            assert abs(r - 1) < (0.10 if lo == 0 else 0.04), (lo, hi, r)
    # the continuum spectrum T / A (no dt factor) misses the top shell by far
    top = live & (kmag >= edges[-2])
    assert s[top].mean() / (T / A[top]).mean() > 1.1  # 1.14 and 1.23 for these cases
    # frozen modes (0 and pi along each axis) receive no noise beyond round-off
    assert np.max(s[~live]) < (1e-20 if precision == 64 else 1e-9)
