"""CPU tests of the Python facade's run_for_time and plot_field (no GPU needed).

run_for_time is driven on a stand-in object that records the run(n) calls, so the
chunking and interrupt semantics of the reference's run_for_time
(activemodelbplus/integrator/__init__.py:105-161) are checked without a device:
  * chunks are exactly step_iterator's sizes, in order;
  * f_interrupt fires before a chunk whenever the elapsed time is an interrupt time,
    and once more after the last chunk;
  * interrupts outside [0, t] and a t that is not a multiple of dt raise ValueError.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "active-model-bplus_amd"))

from activemodelbplus.integrator import Integrator, Stencil, plot_field, step_iterator  # noqa: E402


class FakeSim:
    def __init__(self, dt, timestep=0):
        self.stencil = Stencil(dt, 1.0, 1.0)
        self.timestep = timestep
        self.calls = []

    @property
    def time(self):
        return self.timestep * self.stencil.dt

    @property
    def field(self):
        return np.arange(12.0).reshape(3, 4) * (1 + self.timestep)

    def run(self, n):
        assert isinstance(n, int) and n > 0
        self.calls.append(n)
        self.timestep += n


def test_chunks_follow_step_iterator():
    sim = FakeSim(0.01, timestep=7)
    Integrator.run_for_time(sim, 3.0, max_updates=5, min_update_time=0.5)
    assert sim.calls == list(step_iterator(0.01, 3.0, 5, 0.5))
    assert sum(sim.calls) == 300 and sim.timestep == 307


def test_interrupts_fire_at_their_times():
    sim = FakeSim(0.5)
    seen = []
    Integrator.run_for_time(sim, 5.0, t_interrupt=[0.0, 1.0, 2.5, 5.0], f_interrupt=lambda s: seen.append(s.time),
                            max_updates=2, min_update_time=1.0)
    assert seen == [0.0, 1.0, 2.5, 5.0]
    # the cuts include every interrupt step and the update boundaries (0, 5, 10)
    assert np.cumsum(sim.calls).tolist() == [2, 5, 10]


def test_interrupt_only_at_the_end():
    sim = FakeSim(0.25, timestep=4)  # interrupt times are relative to the start
    seen = []
    Integrator.run_for_time(sim, 2.0, t_interrupt=2.0, f_interrupt=lambda s: seen.append(s.timestep))
    assert seen == [12]


def test_errors():
    sim = FakeSim(0.1)
    with pytest.raises(ValueError):
        Integrator.run_for_time(sim, 0.25)
    with pytest.raises(ValueError):
        Integrator.run_for_time(sim, 1.0, t_interrupt=[1.5], f_interrupt=lambda s: None)
    with pytest.raises(ValueError):
        Integrator.run_for_time(sim, 1.0, t_interrupt=[-0.1], f_interrupt=lambda s: None)
    assert sim.calls == []


def test_plot_field_creates_then_updates():
    mpl = pytest.importorskip("matplotlib")
    mpl.use("Agg")
    import matplotlib.pyplot as plt

    f0 = np.linspace(-1, 1, 20).reshape(4, 5)
    fig, axes = plot_field(f0)
    ax, image, bar = axes
    assert image.get_clim() == (-1.0, 1.0)
    f1 = 3 * f0 + 1
    fig2, axes2 = plot_field(f1, fig, axes)
    assert fig2 is fig and axes2 is axes
    np.testing.assert_array_equal(np.asarray(image.get_array()), f1)
    assert image.get_clim() == (-2.0, 4.0)
    plt.close(fig)
