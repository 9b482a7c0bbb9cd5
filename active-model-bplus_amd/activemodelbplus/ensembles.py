"""Seeded synthetic ensembles of independent Active Model B+ realisations.

``random_ensemble`` draws the heterogeneous ensemble that the batched solver is
benchmarked and tested on: square periodic grids with sides drawn uniformly from
{64, 128, 256, 512}, lambda ~ U[-2, 2], zeta ~ U[-5, 0], phi0 ~ U[-0.6, 0.6],
T in {0, 0.05, 0.1}, a = -0.25, b = 0, c = 0.25, kappa = 1, dt = 5e-3, dx = dy = 1,
phi = phi0 + 0.01 N(0, 1), and noise seed (4 << 32) + i for member i.  No reference
counterpart (the reference integrates one field per Integrator).
"""
from __future__ import annotations

import numpy as np

SIDES = (64, 128, 256, 512)


def random_ensemble(count=512, seed=4, sides=SIDES):
    """Returns (fields, models, seeds, stencil, params) for `count` members; params is
    a dict of the per-member arrays (side, lamb, zeta, phi0, T)."""
    from .integrator import Model, Stencil

    rng = np.random.default_rng(seed)
    side = rng.choice(list(sides), size=count)
    lamb = rng.uniform(-2, 2, size=count)
    zeta = rng.uniform(-5, 0, size=count)
    phi0 = rng.uniform(-0.6, 0.6, size=count)
    temp = rng.choice([0.0, 0.05, 0.1], size=count)
    fields = [phi0[i] + 0.01 * rng.normal(size=(side[i], side[i])) for i in range(count)]
    models = [Model(-0.25, 0, 0.25, 1, lamb[i], zeta[i], temp[i]) for i in range(count)]
    seeds = [(4 << 32) + i for i in range(count)]
    params = {"side": side, "lamb": lamb, "zeta": zeta, "phi0": phi0, "T": temp}
    return fields, models, seeds, Stencil(5e-3, 1, 1), params
