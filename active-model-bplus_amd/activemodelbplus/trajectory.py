"""Trajectories and the structure factor.

The reference does this in examples/structureFactor.ipynb: cell 1 samples a run with
``run_for_time(..., t_interrupt=..., f_interrupt=...)`` and saves
``np.savez(path, time=..., field=..., **model.as_dict(), **stencil.as_dict())``;
cell 7 bins |FFT(phi - phi0) dx dy|^2 by |k| with ``np.digitize(K, bin_edges,
right=True)`` over ``bin_edges = np.linspace(0, 3, 300)[1:]``, takes the bin centres
``k = (bin_edges[1:] + bin_edges[:-1]) / 2``, averages bin ``l`` over the modes with
``indices == l`` and reports <k> = simpson(k S, k) / simpson(S, k) and <R> = 2 pi / <k>.

``structure_factor`` keeps that convention.  Given an ``Integrator`` it bins on the
device (``Integrator.structure_factor_bins``, hipFFT + a binning kernel); given a
host array it does the same arithmetic with numpy (analysis of saved fields).
"""
from __future__ import annotations

import numpy as np

BIN_EDGES = np.linspace(0, 3, 300)[1:]


def save_trajectory(path, times, fields, model, stencil):
    """Writes the reference notebook's trajectory format (cell 1)."""
    params = model.as_dict()
    params.update(stencil.as_dict())
    np.savez(path, time=np.asarray(times, dtype=np.float64), field=np.asarray(fields, dtype=np.float64), **params)


def load_trajectory(path):
    """Reads a trajectory file as a dict (no pickled objects are loaded)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def record(sim, t, times):
    """Runs ``sim`` for time ``t`` and returns (times, fields) sampled at ``times``
    (multiples of dt within [0, t]), as the notebook's interrupt callback does."""
    out_t, out_f = [], []

    def grab(s):
        out_t.append(s.time)
        out_f.append(s.field.copy())

    sim.run_for_time(t, t_interrupt=times, f_interrupt=grab)
    return np.asarray(out_t), np.asarray(out_f)


def structure_factor_bins(field, edges=BIN_EDGES, dx=1.0, dy=1.0):
    """(sums, counts) of |phi_k|^2 = |FFT(phi - mean) dx dy|^2 per index
    ``numpy.digitize(|k|, edges, right=True)`` (len(edges) + 1 entries).  ``field`` is
    an ``Integrator`` (binned on the device, its own dx, dy) or a 2-D host array."""
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    if hasattr(field, "structure_factor_bins"):
        return field.structure_factor_bins(edges)
    phi = np.asarray(field, dtype=np.float64)
    ny, nx = phi.shape
    phiq = np.abs(np.fft.fft2(phi - phi.mean()) * dx * dy) ** 2
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)
    K = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    idx = np.digitize(K, edges, right=True).ravel()
    n = len(edges) + 1
    return np.bincount(idx, weights=phiq.ravel(), minlength=n), np.bincount(idx, minlength=n).astype(np.int64)


def structure_factor(field, edges=BIN_EDGES, dx=1.0, dy=1.0):
    """(k, S) in the notebook's convention: k = bin centres, S[l] = mean of |phi_k|^2
    over the modes with digitize index l (NaN for an empty bin)."""
    sums, counts = structure_factor_bins(field, edges, dx, dy)
    k = 0.5 * (edges[1:] + edges[:-1])
    m = len(k)
    with np.errstate(invalid="ignore", divide="ignore"):
        S = sums[:m] / counts[:m]
    return k, S


def mean_wavenumber(field, edges=BIN_EDGES, dx=1.0, dy=1.0):
    """<k> = simpson(k S, k) / simpson(S, k) over the non-empty bins (the notebook
    integrates all bins; at grids coarser than its 512^2 some bins are empty)."""
    from scipy.integrate import simpson

    k, S = structure_factor(field, edges, dx, dy)
    ok = np.isfinite(S)
    return float(simpson(k[ok] * S[ok], x=k[ok]) / simpson(S[ok], x=k[ok]))


def peak_wavenumber(field, edges=BIN_EDGES, dx=1.0, dy=1.0):
    k, S = structure_factor(field, edges, dx, dy)
    return float(k[np.nanargmax(S)])


def smoothed_peak_wavenumber(field, width=None, dx=1.0, dy=1.0):
    """Wavenumber of the maximum of the structure factor smoothed over |k|: every mode's
    |phi_k|^2 (mean removed) is spread over |k| with a Gaussian of standard deviation
    ``width`` (default: the mode spacing 2 pi / (n dx) of the grid's coarser axis), the
    spread power is divided by the spread mode count, both with weight 1/|k| per mode (the
    modes' density grows as |k|; without the weight the maximum of a ring of modes would sit
    width^2/k0 inside it), and the maximum is located on a grid of width / 50 and refined by
    a parabola through its three points.

    One realisation's bin-averaged S(k) scatters from bin to bin (a bin holds a few tens of
    modes), so its argmax moves by whole bins; the smoothed curve has one maximum whose
    position is stable to a few per cent between realisations.  ``field`` is a 2-D host array."""
    phi = np.asarray(field, dtype=np.float64)
    ny, nx = phi.shape
    p = (np.abs(np.fft.fft2(phi - phi.mean()) * dx * dy) ** 2).ravel()
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=dy)
    K = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2).ravel()
    if width is None:
        width = max(2 * np.pi / (nx * dx), 2 * np.pi / (ny * dy))
    h = width / 50
    keep = K > 0
    idx = np.rint(K[keep] / h).astype(np.int64)
    t = np.arange(-300, 301) * h
    n = max(int(idx.max()) + 1 + 300, t.size)  # >= the kernel's length, so "same" keeps n bins
    inv = 1.0 / K[keep]
    sums = np.bincount(idx, weights=p[keep] * inv, minlength=n)
    counts = np.bincount(idx, weights=inv, minlength=n)
    g = np.exp(-0.5 * (t / width) ** 2)
    num = np.convolve(sums, g, mode="same")
    den = np.convolve(counts, g, mode="same")
    with np.errstate(invalid="ignore", divide="ignore"):
        S = np.where(den > 1e-3 * den.max(), num / den, 0.0)
    i = int(np.argmax(S))
    k = i * h
    if 0 < i < S.size - 1:
        a, b, c = S[i - 1], S[i], S[i + 1]
        d = a - 2 * b + c
        if d < 0:
            k += 0.5 * h * (a - c) / d
    return float(k)


def domain_size(field, edges=BIN_EDGES, dx=1.0, dy=1.0):
    """<R> = 2 pi / <k>, the length the notebook tabulates."""
    return 2 * np.pi / mean_wavenumber(field, edges, dx, dy)
