"""The reference's example scripts (examples/simulateSpinodal.py,
examples/simulateReverseOstwald.py) run unchanged on this package, apart from
show_progress (its live plot needs IPython, which this image lacks): the same
Model/Stencil/Integrator calls, grid, time step and run_for_time chunking.  The spinodal
run is the script's full length (t = 10^3, 10^5 steps); the bubbly-phase run is cut from
t = 10^4 to t = 200."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_simulate_spinodal_example():
    from activemodelbplus import trajectory as T
    from activemodelbplus.integrator import Integrator, Model, Stencil

    stencil = Stencil(dt=1e-2, dx=1, dy=1)
    model = Model(a=-0.25, b=0, c=0.25, kappa=1, lamb=0, zeta=0, T=0)
    phi0 = 0.
    Nx, Ny = 256, 256
    initial = 1e-1 * np.random.default_rng(0).random((Ny, Nx))
    initial += phi0 - np.average(initial)
    tfinal = 1e3
    sim = Integrator(initial, stencil, model)
    k_start = T.mean_wavenumber(sim)
    sim.run_for_time(tfinal, show_progress=False, max_updates=10)
    assert sim.timestep == 100_000 and sim.time == pytest.approx(tfinal)
    field = sim.field
    assert abs(field.sum() - initial.sum()) < 1e-8
    assert 0.8 < np.abs(field).max() < 1.2           # domains near the binodals +-1
    assert T.mean_wavenumber(sim) < 0.5 * k_start     # coarsened


def test_simulate_reverse_ostwald_example():
    from activemodelbplus.integrator import Integrator, Model, Stencil

    stencil = Stencil(dt=1e-2, dx=1, dy=1)
    model = Model(a=-0.25, b=0, c=0.25, kappa=1, lamb=-1, zeta=-4, T=1)
    phi0 = -0.4
    Nx, Ny = 256, 256
    initial = 1e-1 * np.random.default_rng(1).random((Ny, Nx))
    initial += phi0 - np.average(initial)
    tfinal = 200
    sim = Integrator(initial, stencil, model, seed=3)
    sim.run_for_time(tfinal, show_progress=False, max_updates=100)
    assert sim.timestep == 20_000
    s, bad = sim.mass()
    assert bad == 0 and abs(s - initial.sum()) < 1e-6 * initial.size
