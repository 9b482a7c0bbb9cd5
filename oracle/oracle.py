"""ctypes front end of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this module; it is the checker, never the product.  The C restatement it
loads is ``oracle/amb_oracle.c`` (built by ``oracle/Makefile`` into
``oracle/liboracle.so``), which follows test/tjhung_ambplus.h:182-251 and
src/integrator.cu:34-37,96-109,171-175,259-261 of the reference.

Conventions: fields are float64 arrays of shape (ny, nx), row-major; currents are
returned as ``[y, x]`` pairs of (ny, nx) arrays (component 0 = rows, scaled by 1/dy).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class _Stencil(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


class _Model(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("a", "b", "c", "kappa", "lamb", "zeta", "T")]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    src = os.path.join(_HERE, "amb_oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        if force and os.path.exists(_LIB_PATH):
            os.remove(_LIB_PATH)
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.amo_philox4x32_10.argtypes = [u32p, u32p, u32p]
        L.amo_noise_group.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int,
                                      ctypes.c_int64, ctypes.c_int64, dp]
        L.amo_noise.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int64, dp]
        L.amo_terms.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_Stencil),
                                ctypes.POINTER(_Model), dp, dp, dp, dp, dp, dp, dp, dp]
        L.amo_terms.restype = ctypes.c_int
        L.amo_run.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_Stencil),
                              ctypes.POINTER(_Model), ctypes.c_int64, ctypes.c_uint64,
                              ctypes.c_int64, dp, ctypes.c_int]
        L.amo_run.restype = ctypes.c_int
        L.amo_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _stencil(dt, dx, dy):
    return _Stencil(float(dt), float(dx), float(dy))


def _model(a, b, c, kappa, lamb, zeta, T):
    return _Model(float(a), float(b), float(c), float(kappa), float(lamb), float(zeta), float(T))


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (ctypes.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (ctypes.c_uint32 * 4)()
    lib().amo_philox4x32_10(c, k, o)
    return tuple(int(x) for x in o)


def noise_group(seed, step, comp, row, group):
    z = np.zeros(4)
    lib().amo_noise_group(int(seed), int(step), int(comp), int(row), int(group), _dp(z))
    return z


def noise(ny, nx, seed, step):
    """Unscaled normals of one step, shape (2, ny, nx): [xi_y, xi_x]."""
    xi = np.zeros((2, ny, nx))
    lib().amo_noise(int(ny), int(nx), int(seed), int(step), _dp(xi))
    return xi


def terms(phi, stencil, model, xi=None):
    """Every intermediate of one step: dict with mup, mua, jp, jl, jz, jn, divj."""
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    ny, nx = phi.shape
    out = {"mup": np.zeros((ny, nx)), "mua": np.zeros((ny, nx)), "divj": np.zeros((ny, nx))}
    for k in ("jp", "jl", "jz", "jn"):
        out[k] = np.zeros((2, ny, nx))
    if xi is not None:
        xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(2, ny, nx)
    st, md = _stencil(*stencil), _model(*model)
    rc = lib().amo_terms(_dp(phi), ny, nx, ctypes.byref(st), ctypes.byref(md), _dp(xi),
                         _dp(out["mup"]), _dp(out["mua"]), _dp(out["jp"]), _dp(out["jl"]),
                         _dp(out["jz"]), _dp(out["jn"]), _dp(out["divj"]))
    if rc != 0:
        raise ValueError("oracle amo_terms failed")
    return out


def run(phi, stencil, model, nsteps, seed=0, step0=0, noise_buf=None, nthreads=1):
    """nsteps explicit-Euler steps of a copy of phi; returns the new field."""
    phi = np.array(phi, dtype=np.float64, order="C", copy=True)
    ny, nx = phi.shape
    if noise_buf is not None:
        noise_buf = np.ascontiguousarray(noise_buf, dtype=np.float64)
        assert noise_buf.size == nsteps * 2 * ny * nx
    st, md = _stencil(*stencil), _model(*model)
    rc = lib().amo_run(_dp(phi), ny, nx, ctypes.byref(st), ctypes.byref(md), int(nsteps),
                       int(seed), int(step0), _dp(noise_buf), int(nthreads))
    if rc != 0:
        raise ValueError("oracle amo_run failed")
    return phi


def max_threads():
    return int(lib().amo_max_threads())
