"""CPU checks of the kernels' FP64 noise arithmetic (amb_math.h built for the host).

The FP64 step draws its normals with its own logarithm and sine/cosine (no general
library calls; see ln_unit and sincos_turn in amb_math.h).  Pinned here:
  * ln_unit against libm's log on the whole input range u = (w + 1/2) 2^-32, including both
    ends and the sqrt(1/2) reduction boundary: <= 2 ulp of the result;
  * sincos_turn against the exact values at every multiple of 1/8 turn and against libm's
    sin and cos of 2 pi u (which itself rounds 2 pi u): <= 2e-16 absolute;
  * complete normals against the oracle's (libm Box-Muller) for the same Philox counters:
    <= 4e-15 absolute on |z| <= 6.8, far inside what a 1e-10 trajectory bar can see.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "active-model-bplus_amd", "csrc")


@pytest.fixture(scope="module")
def stub(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("noise") / "libnoisestub.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", f"-I{CSRC}",
                    os.path.join(HERE, "native", "noise_math_stub.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    L.stub_normals.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, dp]
    L.stub_box_muller.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_int64, dp]
    L.stub_ln_unit.argtypes = [ctypes.c_double]
    L.stub_ln_unit.restype = ctypes.c_double
    L.stub_sincos_turn.argtypes = [ctypes.c_double, dp, dp]
    return L


def test_ln_unit(stub):
    rng = np.random.default_rng(0)
    words = np.concatenate([np.array([0, 1, 2, 3, 0x7FFFFFFF, 0x80000000, 0xB504F333, 0xB504F334, 0xB504F332,
                                      0xFFFFFFFE, 0xFFFFFFFF], dtype=np.uint64),
                            rng.integers(0, 2**32, size=20000, dtype=np.uint64),
                            2 ** rng.integers(0, 32, size=200).astype(np.uint64)])
    worst = 0.0
    for w in words:
        u = (float(w) + 0.5) * 2.0 ** -32
        ref = math.log(u)
        got = stub.stub_ln_unit(u)
        worst = max(worst, abs(got - ref) / math.ulp(ref))
    assert worst <= 2.0, worst


def test_sincos_turn(stub):
    s, c = ctypes.c_double(), ctypes.c_double()
    exact = {0: (0, 1), 1: (math.sqrt(0.5), math.sqrt(0.5)), 2: (1, 0), 3: (math.sqrt(0.5), -math.sqrt(0.5)),
             4: (0, -1), 5: (-math.sqrt(0.5), -math.sqrt(0.5)), 6: (-1, 0), 7: (-math.sqrt(0.5), math.sqrt(0.5))}
    for k, (es, ec) in exact.items():
        stub.stub_sincos_turn(k / 8, ctypes.byref(s), ctypes.byref(c))
        assert abs(s.value - es) <= 2e-16 and abs(c.value - ec) <= 2e-16, (k, s.value, c.value)
    rng = np.random.default_rng(1)
    worst = 0.0
    for w in np.concatenate([rng.integers(0, 2**32, size=20000), [1, 2**29 - 1, 2**29 + 1, 2**31 - 1, 2**32 - 1]]):
        u = float(w) * 2.0 ** -32
        stub.stub_sincos_turn(u, ctypes.byref(s), ctypes.byref(c))
        th = 2 * math.pi * u
        worst = max(worst, abs(s.value - math.sin(th)), abs(c.value - math.cos(th)))
        assert abs(s.value ** 2 + c.value ** 2 - 1) <= 4e-16
    assert worst <= 2e-15, worst  # libm's reference rounds 2 pi u first (up to 4.4e-16 of angle)


def test_normals_match_oracle(stub, oracle):
    rng = np.random.default_rng(2)
    worst = 0.0
    z = np.zeros(4)
    for _ in range(3000):
        seed, step = int(rng.integers(0, 2**63)), int(rng.integers(0, 2**40))
        comp, row, group = int(rng.integers(0, 2)), int(rng.integers(0, 2**20)), int(rng.integers(0, 2**20))
        stub.stub_normals(seed, step, comp, row, group, z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        ref = oracle.noise_group(seed, step, comp, row, group)
        worst = max(worst, float(np.max(np.abs(z - ref))))
    assert worst <= 4e-15, worst


def test_extreme_words(stub):
    """The smallest u1 gives the largest radius sqrt(-2 ln 2^-33) = 6.769; u2 = 0 gives
    (r, 0) exactly."""
    words = np.array([0, 0, 0xFFFFFFFF, 0x40000000], dtype=np.uint32)
    z = np.zeros(4)
    stub.stub_box_muller(words.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 2,
                         z.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    r0 = math.sqrt(-2 * math.log(2.0 ** -33))
    assert abs(z[0] - r0) <= 4e-15 and z[1] == 0.0
    r1 = math.sqrt(-2 * math.log((2**32 - 0.5) * 2.0 ** -32))
    assert abs(z[2]) <= 1e-20 and abs(z[3] - r1) <= 1e-15 * r1
