"""Planar coexistence: stripes relaxed by the integrator (periodic, 8 rows, FP64, T = 0, five
models in one batched ensemble) against the mean-field binodals of
activemodelbplus.activedroplets.  The passive model coexists at exactly +-1 by symmetry.
With activity the lattice operators (second-order G3 and L3 stencils) move the binodals
from the continuum ones by O(dx^2): at dx = 1 (interface width 2.8 sites) by up to 4e-3,
and halving dx must shrink every deviation about fourfold, so the integrator converges to
the pseudodensity construction (only zeta - 2 lambda sets planar coexistence in the
continuum; on the lattice (-1, 0.5) and (-4, -1) differ at O(dx^2))."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NY = 8
CASES = [(0.0, 0.0), (-1.0, 0.5), (-4.0, -1.0), (1.0, 0.0), (0.0, -1.0)]


def initial_stripe(m, nx, dx):
    """Dense stripe on the middle half of the box with the mean-field planar profile at both
    interfaces (x in physical units)."""
    x = (np.arange(nx) + 0.5) * dx
    L = nx * dx
    prof = m.binodal()
    row = np.where(x < L / 2, prof(L / 4 - x), prof(x - 3 * L / 4))
    return np.tile(row, (NY, 1))


def relax(nx, dx, dt, steps):
    """Plateau densities (dense centre, dilute centre) of every case after ``steps`` steps,
    and their change over a further 2% of the run."""
    import activemodelbplus.integrator as amb
    from activemodelbplus import activedroplets as AD

    theories = [AD.ActiveModelBPlus(zeta, lamb) for zeta, lamb in CASES]
    fields = [initial_stripe(m, nx, dx) for m in theories]
    models = [amb.Model(-0.25, 0, 0.25, 1, lamb, zeta, 0.0) for zeta, lamb in CASES]
    ens = amb.Ensemble(fields, amb.Stencil(dt, dx, dx), models, list(range(len(CASES))), precision=64)
    ens.run(steps)
    before = [ens.field(i) for i in range(len(CASES))]
    ens.run(steps // 50)
    out = []
    for i, m in enumerate(theories):
        f = ens.field(i)
        assert np.all(np.isfinite(f))
        assert np.ptp(f, axis=0).max() < 1e-10  # stays planar
        prof = f.mean(axis=0)
        sim = np.array([prof[nx // 2 - 1 : nx // 2 + 1].mean(), prof[[0, nx - 1]].mean()])
        out.append((np.sort(m.binodal_densities)[::-1], sim, np.abs(f - before[i]).max()))
    return out


def test_planar_coexistence_matches_binodals():
    coarse = relax(96, 1.0, 5e-3, 2_000_000)   # t = 10^4, box 96, dense stripe 48 wide
    fine = relax(192, 0.5, 1e-3, 4_000_000)    # t = 4000, the same box at twice the resolution
    failures = []
    for (zeta, lamb), (th, c, dc), (_, f, df) in zip(CASES, coarse, fine):
        ec, ef = c - th, f - th
        print(f"zeta={zeta:g} lamb={lamb:g}: binodals {th[0]:.6f} {th[1]:.6f}; dx=1: diff {ec[0]:+.2e} {ec[1]:+.2e} "
              f"(drift {dc:.1e}); dx=0.5: diff {ef[0]:+.2e} {ef[1]:+.2e} (drift {df:.1e})", flush=True)
        if max(dc, df) > 2e-7:
            failures.append(("not relaxed", zeta, lamb, dc, df))
        if zeta == 2 * lamb:
            if np.abs(ec).max() > 1e-6 or np.abs(ef).max() > 1e-6:
                failures.append(("passive binodal", zeta, lamb, ec, ef))
        else:
            if np.abs(ec).max() > 5e-3 or np.any(np.abs(ef) > np.abs(ec) / 3):
                failures.append(("active binodal", zeta, lamb, ec, ef))
            if abs(c[0] - 1.0) < 0.05:
                failures.append(("activity did not move the binodal", zeta, lamb, c))
    assert not failures, failures
