"""amb_run_noalloc (the allocation-free run) against amb_run, both through the C ABI.

Two calls in a row on one GPU, the second on a smaller field (the shared static
scratch keeps the first call's contents and must not leak into the second), in both
precisions, with noise: the field equals amb_run's bit for bit and the mass and
non-finite count equal amb_last_mass's.  The test's device buffers come from hipMalloc of
the HIP runtime libamb.so itself links (a second runtime in the process, such as the one
bundled with torch, would not see the same device state).
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "active-model-bplus_amd", "lib", "libamb.so")


class Stencil(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


class Model(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("a", "b", "c", "kappa", "lamb", "zeta", "T")]


class Options(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("seed", ctypes.c_uint64), ("timestep", ctypes.c_int64),
                ("tile", ctypes.c_int)]


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libamb.so is linked against."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7")  # already loaded by libamb.so: the same instance
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def upload(self, dev, host):
        assert self.rt.hipMemcpy(dev, host.ctypes.data_as(ctypes.c_void_p), host.nbytes, 1) == 0  # host to device

    def download(self, host, dev):
        assert self.rt.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), dev, host.nbytes, 2) == 0  # device to host

    def free(self, p):
        assert self.rt.hipFree(p) == 0


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(LIB)
    dp = ctypes.POINTER(ctypes.c_double)
    L.amb_create.argtypes = [dp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(Stencil), ctypes.POINTER(Model),
                             ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_void_p)]
    L.amb_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
    L.amb_get_field.argtypes = [ctypes.c_void_p, dp]
    L.amb_last_mass.argtypes = [ctypes.c_void_p, dp, ctypes.POINTER(ctypes.c_int64)]
    L.amb_destroy.argtypes = [ctypes.c_void_p]
    L.amb_last_error.restype = ctypes.c_char_p
    L.amb_run_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                  ctypes.POINTER(Stencil), ctypes.POINTER(Model), ctypes.c_uint64, ctypes.c_int64,
                                  ctypes.c_int64, ctypes.c_void_p, dp, ctypes.POINTER(ctypes.c_int64)]
    return L


def with_handle(lib, phi, st, md, precision, seed, timestep, nsteps):
    h = ctypes.c_void_p()
    opt = Options(precision, seed, timestep, 0)
    a = np.ascontiguousarray(phi, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    assert lib.amb_create(a.ctypes.data_as(dp), a.shape[0], a.shape[1], ctypes.byref(st), ctypes.byref(md),
                          ctypes.byref(opt), ctypes.byref(h)) == 0, lib.amb_last_error()
    try:
        assert lib.amb_run(h, nsteps) == 0, lib.amb_last_error()
        out = np.empty_like(a)
        assert lib.amb_get_field(h, out.ctypes.data_as(dp)) == 0
        s, bad = ctypes.c_double(), ctypes.c_int64()
        assert lib.amb_last_mass(h, ctypes.byref(s), ctypes.byref(bad)) == 0
        return out, s.value, bad.value
    finally:
        lib.amb_destroy(h)


def without_alloc(lib, phi, st, md, precision, seed, timestep, nsteps):
    hip = Hip()
    host = np.ascontiguousarray(phi, dtype=np.float32 if precision == 32 else np.float64)
    field, work = hip.alloc(host.nbytes), hip.alloc(host.nbytes)
    try:
        hip.upload(field, host)
        hip.upload(work, np.full_like(host, 123.0))  # arbitrary content: overwritten, never read
        s, bad = ctypes.c_double(), ctypes.c_int64()
        rc = lib.amb_run_noalloc(field, work, phi.shape[0], phi.shape[1], precision, ctypes.byref(st),
                                 ctypes.byref(md), seed, timestep, nsteps, None, ctypes.byref(s), ctypes.byref(bad))
        assert rc == 0, lib.amb_last_error()
        out = np.empty_like(host)
        hip.download(out, field)
        return out.astype(np.float64), s.value, bad.value
    finally:
        hip.free(field)
        hip.free(work)


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_matches_amb_run_twice(lib, precision):
    st = Stencil(5e-3, 1.0, 1.0)
    md = Model(-0.25, 0.0, 0.25, 1.0, -1.0, -4.0, 0.1)
    rng = np.random.default_rng(precision)
    cases = [((384, 320), 77, 5, 45), ((96, 136), 78, 0, 33)]  # the second call on a smaller field
    for shape, seed, t0, nsteps in cases:
        phi = -0.4 + 0.01 * rng.normal(size=shape)
        print(f"noalloc FP{precision} {shape} seed {seed} t0 {t0} steps {nsteps}", flush=True)
        f_ref, s_ref, b_ref = with_handle(lib, phi, st, md, precision, seed, t0, nsteps)
        f, s, b = without_alloc(lib, phi, st, md, precision, seed, t0, nsteps)
        np.testing.assert_array_equal(f, f_ref)
        assert b == b_ref == 0
        assert s == s_ref


def test_noalloc_rejects_large_fields_and_bad_args(lib):
    hip = Hip()
    st = Stencil(5e-3, 1.0, 1.0)
    md = Model(-0.25, 0.0, 0.25, 1.0, 0.0, 0.0, 0.0)
    a, b = hip.alloc(64), hip.alloc(64)
    try:
        # 16384^2 FP32 has more tiles than the scratch area holds: rejected before any launch
        rc = lib.amb_run_noalloc(a, b, 16384, 16384, 32, ctypes.byref(st), ctypes.byref(md), 0, 0, 1, None, None,
                                 None)
        assert rc == 1 and b"too large" in lib.amb_last_error()
        rc = lib.amb_run_noalloc(a, a, 2, 2, 32, ctypes.byref(st), ctypes.byref(md), 0, 0, 1, None, None, None)
        assert rc == 1
    finally:
        hip.free(a)
        hip.free(b)
