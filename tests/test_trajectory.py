"""The trajectory format and the structure factor of activemodelbplus.trajectory
against the reference's examples/structureFactor.ipynb: cell 1's file layout and
cell 7's binning (a literal loop over bins, as the notebook writes it), a single
Fourier mode (all its power in one bin, with the closed-form amplitude), and
Parseval's identity over all digitize indices."""
import numpy as np
import pytest

from activemodelbplus import trajectory as T
from activemodelbplus.integrator import Model, Stencil


def notebook_S(phi, dx, dy, bin_edges):
    """Cell 7's loop: S[l] = average of |phi_q|^2 over indices == l."""
    ny, nx = phi.shape
    phiq = np.abs(np.fft.fftshift(np.fft.fft2(phi - np.average(phi))) * dx * dy) ** 2
    kx = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx, d=dx))
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny, d=dy))
    Kx, Ky = np.meshgrid(kx, ky, indexing="xy")
    indices = np.digitize((Kx**2 + Ky**2) ** 0.5, bin_edges, right=True)
    k = 0.5 * (bin_edges[1:] + bin_edges[:-1])
    S = np.zeros_like(k)
    for l in range(S.size):
        sel = phiq[indices == l]
        S[l] = sel.mean() if sel.size else np.nan
    return k, S


@pytest.mark.parametrize("shape,dx,dy", [((64, 64), 1.0, 1.0), ((48, 80), 0.75, 1.25), ((33, 17), 1.0, 0.5)])
def test_binning_matches_notebook(shape, dx, dy):
    phi = np.random.default_rng(1).normal(size=shape)
    k, S = T.structure_factor(phi, dx=dx, dy=dy)
    k_nb, S_nb = notebook_S(phi, dx, dy, T.BIN_EDGES)
    np.testing.assert_array_equal(k, k_nb)
    np.testing.assert_array_equal(np.isnan(S), np.isnan(S_nb))
    ok = ~np.isnan(S)
    np.testing.assert_allclose(S[ok], S_nb[ok], rtol=1e-12)


def test_single_mode():
    n, m = 128, 5
    x = np.arange(n)
    phi = np.cos(2 * np.pi * m * x / n)[None, :] * np.ones((n, 1))
    sums, counts = T.structure_factor_bins(phi)
    k0 = 2 * np.pi * m / n
    i = int(np.digitize([k0], T.BIN_EDGES, right=True)[0])
    assert sums[i] == pytest.approx(n**4 / 2, rel=1e-12)  # two modes of |n^2 / 2|^2
    others = np.delete(sums, i)
    assert np.max(others) < 1e-12 * n**4


def test_parseval_over_all_bins():
    phi = np.random.default_rng(2).normal(size=(40, 56))
    dx, dy = 0.5, 2.0
    sums, counts = T.structure_factor_bins(phi, dx=dx, dy=dy)
    assert counts.sum() == phi.size
    want = phi.size * np.sum((phi - phi.mean()) ** 2) * (dx * dy) ** 2
    assert sums.sum() == pytest.approx(want, rel=1e-12)


def test_mean_wavenumber_of_a_ring():
    """A field made of modes on one shell |k| = k0 has <k> within a bin of k0."""
    n = 256
    y, x = np.mgrid[0:n, 0:n]
    phi = np.zeros((n, n))
    for mx, my in [(12, 0), (0, 12), (9, 8), (8, 9)]:
        phi += np.cos(2 * np.pi * (mx * x + my * y) / n)
    k0 = 2 * np.pi * 12 / n
    assert abs(T.peak_wavenumber(phi) - k0) < 0.02
    assert abs(T.mean_wavenumber(phi) - k0) < 0.02
    assert T.domain_size(phi) == pytest.approx(2 * np.pi / T.mean_wavenumber(phi))


def test_trajectory_file_layout(tmp_path):
    fields = np.random.default_rng(3).normal(size=(3, 8, 8))
    path = tmp_path / "trajectory_1.npz"
    T.save_trajectory(path, [1.0, 10.0, 100.0], fields, Model(-0.25, 0, 0.25, 1, 0.5, -1, 0), Stencil(1e-2, 1, 1))
    z = T.load_trajectory(path)
    assert set(z) == {"time", "field", "a", "b", "c", "kappa", "lamb", "zeta", "temperature", "dt", "dx", "dy"}
    np.testing.assert_array_equal(z["field"], fields)
    assert float(z["lamb"]) == 0.5 and float(z["dt"]) == 1e-2


def test_smoothed_peak_of_a_ring_and_its_invariances():
    """Power on the modes of one lattice ring |k| = k0 (plus weaker modes at 2 k0): the
    smoothed maximum lies within a tenth of a mode spacing of k0, also with a kernel twice as
    wide (the 1/|k| weight removes the pull of the growing mode density).  Amplitude scaling,
    a constant offset and a lattice translation leave it unchanged."""
    n = 256
    y, x = np.mgrid[0:n, 0:n]
    m = 5  # k0 = 2 pi 5 / 256 = 0.123: the lattice vectors of length 5
    phi = np.zeros((n, n))
    for a, b in [(5, 0), (0, 5), (3, 4), (4, 3), (3, -4), (4, -3)]:
        phi += np.cos(2 * np.pi * (a * x + b * y) / n + 0.3 * a)
    phi += 0.3 * np.cos(2 * np.pi * 10 * x / n) + 0.3 * np.cos(2 * np.pi * (6 * x + 8 * y) / n)
    k0 = 2 * np.pi * m / n
    w = 2 * np.pi / n
    kp = T.smoothed_peak_wavenumber(phi)
    assert abs(kp - k0) < 0.1 * w, (kp, k0)
    assert T.smoothed_peak_wavenumber(3.7 * phi + 0.2) == pytest.approx(kp, rel=1e-12)
    assert T.smoothed_peak_wavenumber(np.roll(phi, (17, -5), axis=(0, 1))) == pytest.approx(kp, rel=1e-9)
    assert abs(T.smoothed_peak_wavenumber(phi, width=2 * w) - k0) < 0.1 * w


def test_smoothed_peak_matches_a_direct_evaluation():
    """The binned convolution equals the smoothed curve written out mode by mode, on grids
    where the kernel is longer than the spectrum (8x8) and shorter (24x40, anisotropic): the
    same maximum to the fine grid's resolution."""
    rng = np.random.default_rng(3)
    for shape, dx, dy in [((8, 8), 1.0, 1.0), ((24, 40), 0.75, 1.25)]:
        phi = rng.normal(size=shape)
        ny, nx = shape
        p = (np.abs(np.fft.fft2(phi - phi.mean()) * dx * dy) ** 2).ravel()
        K = np.sqrt((2 * np.pi * np.fft.fftfreq(nx, d=dx))[None, :] ** 2
                    + (2 * np.pi * np.fft.fftfreq(ny, d=dy))[:, None] ** 2).ravel()
        keep = K > 0
        width = max(2 * np.pi / (nx * dx), 2 * np.pi / (ny * dy))
        grid = np.arange(0, K.max() + 6 * width, width / 50)
        w = np.exp(-0.5 * ((grid[:, None] - K[keep][None, :]) / width) ** 2) / K[keep][None, :]
        S = (w @ p[keep]) / w.sum(axis=1)
        k_direct = grid[np.argmax(S)]
        assert abs(T.smoothed_peak_wavenumber(phi, dx=dx, dy=dy) - k_direct) < 1.5 * width / 50
