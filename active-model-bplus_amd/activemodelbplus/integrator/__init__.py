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

__all__ = ["Model", "Stencil", "Integrator", "Ensemble", "Surrogate", "CudaException", "HipException",
           "steps_from_time", "step_iterator", "plot_field", "device_count", "tile_shapes"]


def plot_field(field, fig=None, axes=None, fontsize=8):
    """Show the evolving field during program execution (reference :14-36)."""
    import matplotlib.pyplot as plt
    from matplotlib.colors import Normalize

    if fig is not None:
        ax, im, cbar = axes
        im.set_data(field)
        im.set_norm(Normalize(field.min(), vmax=field.max()))
        return fig, axes

    fig = plt.figure()
    ax = plt.gca()
    im = ax.imshow(field)
    cbar = fig.colorbar(im, ax=ax, fraction=0.05, pad=2e-2, location='right')
    cbar.set_label(r'$\phi$', fontsize=fontsize)
    cbar.ax.tick_params(labelsize=fontsize)
    ax.set_xlabel('$x$', fontsize=fontsize)
    ax.set_ylabel('$y$', fontsize=fontsize)
    ax.set_xticks([])
    ax.set_yticks([])
    return fig, (ax, im, cbar)


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


class Integrator(BaseIntegrator):
    def run_for_time(self, t, show_progress=False, t_interrupt=None, f_interrupt=None,
                     max_updates=100, min_update_time=10):
        """Run the simulation for a span of time (reference :105-161).

        Args:
            t: time to run for; a multiple of the stencil's dt.
            show_progress: show a live plot and a progress bar (needs matplotlib,
                tqdm and IPython).
            t_interrupt: times (multiples of dt, within [0, t]) at which
                ``f_interrupt(self)`` is called.
        Raises:
            ValueError: if t is not a multiple of the timestep, or an interrupt
                lies outside [0, t].
        """
        nsteps = steps_from_time(self.stencil.dt, t)

        if t_interrupt is not None:
            assert f_interrupt is not None
            if np.any(np.asarray(t_interrupt) < 0) or np.any(np.asarray(t_interrupt) > t):
                raise ValueError('interrupt outside sim time window!')

        initial_timestep = self.timestep
        initial_time = self.time
        if show_progress:
            fig, axes = plot_field(self.field)

        steps = step_iterator(self.stencil.dt, t, max_updates, min_update_time,
                              t_interrupt=t_interrupt)
        if show_progress:
            from tqdm import tqdm
            steps = tqdm(steps)

        for n in steps:
            if show_progress:
                for out in [sys.stdout, sys.stderr]:
                    out.flush()

            if t_interrupt is not None:
                if np.any(np.isclose(self.time - initial_time, t_interrupt)):
                    f_interrupt(self)

            assert n > 0
            self.run(int(n))

            if show_progress:
                from IPython.display import clear_output, display
                clear_output(wait=True)
                plot_field(self.field, fig, axes)
                display(fig)

        if show_progress:
            import matplotlib.pyplot as plt
            plt.close()
        assert (self.timestep - initial_timestep) == nsteps

        if t_interrupt is not None:
            if np.any(np.isclose(self.time - initial_time, t_interrupt)):
                f_interrupt(self)

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
