"""What the trained surrogate (weights shipped in activemodelbplus/data) gains over the
method without the network.  Over the same span, from an Euler state at t = 50 on a
held-out parameter draw (256 x 256, T = 0):
  * twenty surrogate coarse steps of Delta = 100 dt track 2000 explicit-Euler steps
    to a few per cent (L2);
  * explicit Euler with the coarse step itself (dt = Delta = 0.5, the scheme without the
    network) is unstable there and raises (non-finite field);
  * the surrogate is at least twice as close to Euler as persistence (phi unchanged);
  * mass is conserved to round-off."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIDE, DT, DELTA = 256, 5e-3, 100


@pytest.fixture(scope="module")
def trained():
    import activemodelbplus.integrator as amb
    from activemodelbplus import surrogate as SG
    path = SG.default_weights_path()
    assert os.path.exists(path), "trained weights are part of the package"
    S = amb.Surrogate(SIDE, SIDE, max_batch=1, seed=0)
    SG.load_weights(S, path)
    return amb, SG, S


def euler_state(amb, lamb, zeta, phi0, t, seed=21):
    rng = np.random.default_rng(seed)
    phi = phi0 + 0.01 * rng.normal(size=(SIDE, SIDE))
    sim = amb.Integrator(phi, amb.Stencil(DT, 1, 1), amb.Model(-0.25, 0, 0.25, 1, lamb, zeta, 0), seed=0,
                         precision=32)
    sim.run(int(round(t / DT)))
    return sim


@pytest.mark.parametrize("lamb,zeta,phi0", [(0.7, -1.5, 0.0), (-1.2, -3.5, 0.2)])
def test_surrogate_gain_over_coarse_euler(trained, lamb, zeta, phi0):
    amb, SG, S = trained
    sim = euler_state(amb, lamb, zeta, phi0, 50.0)
    start = sim.field
    coarse_steps = 20
    # reference: explicit Euler at dt
    sim.run(coarse_steps * DELTA)
    ref = sim.field
    # surrogate
    sur = amb.Integrator(start, amb.Stencil(DT, 1, 1), amb.Model(-0.25, 0, 0.25, 1, lamb, zeta, 0), seed=0,
                         precision=32)
    S.advance(sur, coarse_steps, [lamb, zeta, float(start.mean())], DELTA)
    got = sur.field
    e_s = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    e_p = np.linalg.norm(start - ref) / np.linalg.norm(ref)
    print(f"surrogate {e_s:.3e}  persistence {e_p:.3e}")
    assert e_s < 0.05
    assert e_s < 0.5 * e_p
    assert abs(got.sum() - start.sum()) < 1e-2
    # the method without the network: Euler with the coarse step
    coarse = amb.Integrator(start, amb.Stencil(DT * DELTA, 1, 1), amb.Model(-0.25, 0, 0.25, 1, lamb, zeta, 0),
                            seed=0, precision=32)
    with pytest.raises(amb.CudaException):
        coarse.run(coarse_steps)


def test_hybrid_reaches_structure_factor_of_euler(trained):
    """Short hybrid rollout (t = 50 -> 150) with audits: <k> within 2% of Euler."""
    amb, SG, S = trained
    lamb, zeta, phi0 = 0.5, -1.0, 0.0
    sim = euler_state(amb, lamb, zeta, phi0, 50.0, seed=22)
    start = sim.field
    sim.run(20000)
    ref = sim.field
    h = SG.HybridIntegrator(S, start, amb.Model(-0.25, 0, 0.25, 1, lamb, zeta, 0), amb.Stencil(DT, 1, 1),
                            audit_every=20)
    h.run_coarse(200)
    k_e, k_h = SG.mean_wavenumber(ref), SG.mean_wavenumber(h.field)
    print(f"<k> euler {k_e:.4f} hybrid {k_h:.4f}  stats {dict((k, v) for k, v in h.stats.items() if k != 'audit_errors')}")
    assert abs(k_h / k_e - 1) < 0.02
    assert h.stats["audits"] == 10
