"""A field of more than 2^31 sites (96 x 22,369,664 FP32, 17 GB as the float64 input):
every index of the fused step, the mass reduction and the host copies must be 64-bit.
The field is a 96 x 64 pattern repeated along x; at T = 0 each site sees the same
periodic neighbourhood as in a plain 96 x 64 run, so after the same steps every copy of
the pattern equals that run bit for bit and the FP64 mass is the copies' sum."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NY, PW, REPS = 96, 64, 349_526  # NY * PW * REPS = 2,147,487,744 > 2^31 sites
STEPS = 20


def test_field_beyond_2_31_sites():
    import activemodelbplus.integrator as amb

    assert NY * PW * REPS > 2 ** 31
    t0 = time.perf_counter()
    rng = np.random.default_rng(31)
    pattern = -0.2 + 0.1 * rng.normal(size=(NY, PW))
    st, md = amb.Stencil(5e-3, 1, 1), amb.Model(-0.25, 0, 0.25, 1, 0.7, -1.5, 0.0)
    small = amb.Integrator(pattern, st, md, seed=3, precision=32)
    small.run(STEPS)
    ref = small.field
    mass_small, _ = small.mass()

    big = amb.Integrator(np.tile(pattern, (1, REPS)), st, md, seed=3, precision=32)
    print(f"created {NY} x {PW * REPS} after {time.perf_counter() - t0:.1f} s", flush=True)
    big.run(STEPS)
    mass_big, bad = big.mass()
    print(f"ran {STEPS} steps after {time.perf_counter() - t0:.1f} s", flush=True)
    assert bad == 0
    f = big.field
    assert f.shape == (NY, PW * REPS)
    print(f"copied back after {time.perf_counter() - t0:.1f} s", flush=True)
    view = f.reshape(NY, REPS, PW)
    assert np.array_equal(view[:, 0, :], ref)
    assert np.array_equal(view[:, -1, :], ref)
    assert bool((view == ref[:, None, :]).all())
    del view, f
    assert mass_big == pytest.approx(REPS * mass_small, rel=1e-9, abs=1e-3)
