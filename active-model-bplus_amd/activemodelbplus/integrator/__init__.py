#!/usr/bin/env python3
"""Python facade of the MI355X Active Model B+ integrator.

Same public names and behaviour as the reference's
activemodelbplus/integrator/__init__.py:1-161: ``Model``, ``Stencil``,
``Integrator`` (with ``run_for_time``), ``steps_from_time``, ``step_iterator``
and ``plot_field``.  The compiled module ``integrator`` calls the HIP kernels
through the C ABI of ``lib/libamb.so``; there is no CPU fallback, so importing
works anywhere but constructing an ``Integrator`` needs an AMD GPU.

Differences from the reference, all additive:
  * matplotlib / tqdm / IPython are imported only when ``show_progress=True``
    (the reference imports them at module load, :7-11);
  * ``Integrator`` takes keyword-only ``seed``, ``precision`` (64 or 32),
    ``timestep`` and ``tile``; with a seed the noise is reproducible and
    ``checkpoint()`` / ``Integrator.from_checkpoint`` resume a run exactly;
  * ``Ensemble`` advances many independent realisations in one launch per step.
"""
import sys

import numpy as np

from .integrator import CudaException, HipException, Model, Stencil
from .integrator import Ensemble as _BaseEnsemble
from .integrator import Integrator as BaseIntegrator
from .integrator import Surrogate, device_count, tile_shapes
from .integrator import tuning_clear, tuning_entries, tuning_get, tuning_set

__all__ = ["Model", "Stencil", "Integrator", "Ensemble", "Surrogate", "CudaException", "HipException",
           "steps_from_time", "step_iterator", "plot_field", "device_count", "tile_shapes",
           "tuning_set", "tuning_get", "tuning_entries", "tuning_clear"]


def plot_field(field, fig=None, axes=None, fontsize=8):
    """Live image of a field with a colour bar.

    The first call (``fig=None``) opens a figure and returns ``(fig, (ax, image,
    colorbar))``; passing those back redraws the same image with the new field and a
    colour range stretched to its minimum and maximum.  Same call pattern as the
    reference's helper (activemodelbplus/integrator/__init__.py:14-36); matplotlib is
    imported only here.
    """
    values = np.asarray(field)
    lo, hi = float(values.min()), float(values.max())
    if fig is None:
        import matplotlib.pyplot as plt

        fig, ax = plt.subplots()
        image = ax.imshow(values, vmin=lo, vmax=hi)
        ax.set_xticks([])
        ax.set_yticks([])
        for name, setter in (("$x$", ax.set_xlabel), ("$y$", ax.set_ylabel)):
            setter(name, fontsize=fontsize)
        colorbar = fig.colorbar(image, ax=ax, fraction=0.05, pad=0.02)
        colorbar.set_label(r"$\phi$", fontsize=fontsize)
        colorbar.ax.tick_params(labelsize=fontsize)
        return fig, (ax, image, colorbar)
    image = axes[1]
    image.set_data(values)
    image.set_clim(lo, hi)
    return fig, axes


def steps_from_time(dt, tfinal):
    """Number of steps that make up an elapsed time (reference :39-57).

    Raises:
        ValueError: if tfinal is not a multiple of dt.
    """
    nsteps = np.asarray(np.round(tfinal / dt), dtype=int)
    error = np.abs(nsteps - tfinal / dt)
    if not np.all(np.isclose(error, 0)):
        raise ValueError(f'tfinal={tfinal} is not a multiple of dt={dt}!')
    return nsteps.item() if np.isscalar(nsteps) or nsteps.ndim == 0 else nsteps


def step_iterator(dt, tfinal, max_updates, min_update_time, t_interrupt=None):
    """Chunk sizes that split a run into roughly even updates (reference :60-102).

    Chunk boundaries are linspace(0, nsteps, nupdates + 1) with
    nupdates = clamp(nsteps // round(min_update_time / dt), 1, max_updates), merged
    with the interrupt steps.
    """
    min_steps = int(np.round(min_update_time / dt))
    min_steps = max(min_steps, 1)

    nsteps = steps_from_time(dt, tfinal)
    nupdates = nsteps // min_steps
    nupdates = min(nupdates, max_updates)
    nupdates = max(nupdates, 1)

    timestep = np.linspace(0, nsteps, nupdates + 1, dtype=int)
    if t_interrupt is not None:
        t_interrupt = np.asarray(t_interrupt).reshape(-1)
        interrupts = steps_from_time(dt, t_interrupt)
        timestep = np.concatenate([timestep, np.asarray(interrupts).reshape(-1)])
        timestep = np.unique(timestep)

    step_size = np.diff(timestep)
    assert len(step_size) >= 1
    assert np.sum(step_size) == nsteps
    assert np.all(step_size >= 0)
    return step_size


class _LiveView:
    """Progress display of run_for_time: a tqdm bar over the chunks and a redrawn image
    of the field after each one (matplotlib, tqdm and IPython are needed only here)."""

    def __init__(self, sim):
        self.sim = sim
        self.fig, self.axes = plot_field(sim.field)

    def chunks(self, sizes):
        from tqdm import tqdm
        return tqdm(sizes)

    def before_chunk(self):
        sys.stdout.flush()
        sys.stderr.flush()

    def after_chunk(self):
        from IPython.display import clear_output, display
        clear_output(wait=True)
        plot_field(self.sim.field, self.fig, self.axes)
        display(self.fig)

    def close(self):
        import matplotlib.pyplot as plt
        plt.close(self.fig)


class Integrator(BaseIntegrator):
    def run_for_time(self, t, show_progress=False, t_interrupt=None, f_interrupt=None,
                     max_updates=100, min_update_time=10):
        """Advance the simulation by a span of time t (a multiple of the stencil's dt).

        The span is cut into chunks by ``step_iterator`` (at most ``max_updates`` of them,
        each at least ``min_update_time`` long, plus a cut at every interrupt time); each
        chunk is one ``run(n)``.  ``f_interrupt(self)`` is called whenever the time elapsed
        since the start matches one of ``t_interrupt``: before a chunk, and once more
        after the last one.  Behaviour of the reference's run_for_time
        (activemodelbplus/integrator/__init__.py:105-161).

        Raises:
            ValueError: if t (or an interrupt time) is not a multiple of dt, or an
                interrupt lies outside [0, t].
        """
        dt = self.stencil.dt
        total = steps_from_time(dt, t)
        samples = None
        if t_interrupt is not None:
            assert f_interrupt is not None
            samples = np.asarray(t_interrupt, dtype=float)
            if (samples < 0).any() or (samples > t).any():
                raise ValueError('interrupt outside sim time window!')
        sizes = step_iterator(dt, t, max_updates, min_update_time, t_interrupt=t_interrupt)

        start_step, start_time = self.timestep, self.time

        def sample_if_due():
            if samples is not None and np.isclose(self.time - start_time, samples).any():
                f_interrupt(self)

        view = _LiveView(self) if show_progress else None
        for n in (view.chunks(sizes) if view else sizes):
            if view:
                view.before_chunk()
            sample_if_due()
            assert n > 0
            self.run(int(n))
            if view:
                view.after_chunk()
        if view:
            view.close()
        assert self.timestep - start_step == total
        sample_if_due()

    # -- exact resume (no reference counterpart; the reference reseeds on every
    #    construction, src/integrator.cu:320-323) --------------------------------
    def checkpoint(self):
        """Everything needed to resume bit-exactly: field, timestep, seed, params."""
        opts = self.options
        state = {"field": self.field, "timestep": self.timestep, "seed": opts["seed"],
                 "precision": opts["precision"]}
        state.update({"model": self.model.as_tuple(), "stencil": self.stencil.as_tuple()})
        return state

    def save(self, path):
        s = self.checkpoint()
        np.savez(path, field=s["field"], timestep=s["timestep"], seed=np.uint64(s["seed"]),
                 precision=s["precision"], model=np.asarray(s["model"]),
                 stencil=np.asarray(s["stencil"]))

    @classmethod
    def from_checkpoint(cls, state, **kwargs):
        if isinstance(state, str):
            with np.load(state, allow_pickle=False) as z:
                state = {k: z[k] for k in z.files}
        model = Model(*[float(x) for x in np.asarray(state["model"]).reshape(-1)])
        stencil = Stencil(*[float(x) for x in np.asarray(state["stencil"]).reshape(-1)])
        kw = dict(seed=int(state["seed"]), precision=int(state["precision"]),
                  timestep=int(state["timestep"]))
        kw.update(kwargs)
        return cls(np.asarray(state["field"]), stencil, model, **kw)


class Ensemble(_BaseEnsemble):
    """Independent realisations advanced together (no reference counterpart).

    ``Ensemble(fields, stencils, models, seeds, precision=32)``; ``stencils`` and
    ``models`` may be single objects shared by every realisation.
    """

    def __init__(self, fields, stencils, models, seeds, precision=32):
        n = len(fields)
        if isinstance(stencils, Stencil):
            stencils = [stencils] * n
        if isinstance(models, Model):
            models = [models] * n
        super().__init__([np.ascontiguousarray(f, dtype=np.float64) for f in fields],
                         list(stencils), list(models), [int(s) for s in seeds], precision)
