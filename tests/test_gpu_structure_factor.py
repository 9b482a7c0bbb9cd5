"""The device structure factor (hipFFT + binning kernel, Integrator.structure_factor_bins)
against the numpy path of activemodelbplus.trajectory, which is pinned to the
reference notebook in tests/test_trajectory.py."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,shape,dx,dy", [(64, (256, 256), 1.0, 1.0), (64, (96, 160), 0.75, 1.25),
                                                   (32, (128, 128), 1.0, 1.0), (32, (64, 200), 1.0, 0.5),
                                                   (64, (60, 75), 1.0, 1.0), (32, (33, 51), 1.0, 1.0)])
def test_device_bins_match_numpy(precision, shape, dx, dy):
    import activemodelbplus.integrator as amb
    from activemodelbplus import trajectory as T
    rng = np.random.default_rng(7)
    phi = 0.3 * rng.normal(size=shape) + 0.1
    sim = amb.Integrator(phi, amb.Stencil(1e-3, dx, dy), amb.Model(-0.25, 0, 0.25, 1, 0, 0, 0), seed=1,
                         precision=precision)
    field = sim.field  # as stored (FP32 rounds the input)
    sums, counts = sim.structure_factor_bins(T.BIN_EDGES)
    ref_sums, ref_counts = T.structure_factor_bins(field, dx=dx, dy=dy)
    np.testing.assert_array_equal(counts, ref_counts)
    tol = 1e-11 if precision == 64 else 2e-5
    scale = ref_sums.max()
    assert np.max(np.abs(sums - ref_sums)) <= tol * scale
    assert T.mean_wavenumber(sim) == pytest.approx(T.mean_wavenumber(field, dx=dx, dy=dy), rel=10 * tol)
