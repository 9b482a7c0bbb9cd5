"""A second, independent CPU restatement of one Active Model B+ step in NumPy
(np.roll periodic shifts), used to cross-check the C oracle on tiny grids.

Written from the stencil definitions of the reference
(src/finite_difference.h:73,379-412; src/finite_difference.cuh:29-39,166-183;
src/integrator.cu:34-37,96-109,171-175,259-261).  Component 0 = rows (y, 1/dy),
component 1 = columns (x, 1/dx).
"""
import numpy as np

D8_TAPS = {1: 4 / 5, 2: -1 / 5, 3: 4 / 105, 4: -1 / 280}


def sh(f, dr, dc):
    """f shifted so that out[i, j] = f[i + dr, j + dc] (periodic)."""
    return np.roll(np.roll(f, -dr, axis=0), -dc, axis=1)


def d8_x(f, dx):
    return sum(w * (sh(f, 0, k) - sh(f, 0, -k)) for k, w in D8_TAPS.items()) / dx


def d8_y(f, dy):
    return sum(w * (sh(f, k, 0) - sh(f, -k, 0)) for k, w in D8_TAPS.items()) / dy


def g3_x(f, dx):
    # rows (0.1, 0.3, 0.1) x columns (-1, 0, 1)
    s = {-1: 0.1, 0: 0.3, 1: 0.1}
    return sum(w * (sh(f, a, 1) - sh(f, a, -1)) for a, w in s.items()) / dx


def g3_y(f, dy):
    s = {-1: 0.1, 0: 0.3, 1: 0.1}
    return sum(w * (sh(f, 1, b) - sh(f, -1, b)) for b, w in s.items()) / dy


def l3(f, dx, dy):
    s = {-1: -0.25, 0: 1.5, 1: -0.25}
    sx = sum(w * (sh(f, a, 1) - 2 * sh(f, a, 0) + sh(f, a, -1)) for a, w in s.items())
    sy = sum(w * (sh(f, 1, b) - 2 * sh(f, 0, b) + sh(f, -1, b)) for b, w in s.items())
    return sx / dx**2 + sy / dy**2


def terms(phi, stencil, model, xi=None):
    dt, dx, dy = stencil
    a, b, c, kappa, lamb, zeta, T = model
    gy, gx, lap = g3_y(phi, dy), g3_x(phi, dx), l3(phi, dx, dy)
    mup = a * phi + b * phi**2 + c * phi**3 - kappa * lap
    mua = lamb * (gy**2 + gx**2)
    jz = np.stack([zeta * lap * gy, zeta * lap * gx])
    amp = np.sqrt(2 * T) * np.sqrt(1 / (dt * dx * dy))
    jn = np.zeros_like(jz) if xi is None else amp * np.asarray(xi).reshape(jz.shape)
    jp = np.stack([-d8_y(mup, dy), -d8_x(mup, dx)])
    jl = np.stack([-g3_y(mua, dy), -g3_x(mua, dx)])
    j1, j2 = jp + jn, jl + jz
    divj = g3_y(j2[0], dy) + g3_x(j2[1], dx) + d8_y(j1[0], dy) + d8_x(j1[1], dx)
    return dict(mup=mup, mua=mua, jp=jp, jl=jl, jz=jz, jn=jn, divj=divj)


def step(phi, stencil, model, xi=None):
    return phi - stencil[0] * terms(phi, stencil, model, xi)["divj"]
