"""Pins the CPU oracle (oracle/amb_oracle.c) before it is trusted as the checker.

Every check here is independent of the oracle's own formulas:
  * Philox4x32-10 against the published Random123 known-answer vectors and
    against rocRAND's open-source implementation (compiled on the host);
  * the reference's own polynomial KATs (test/test_finite_differences.cc:48-121)
    for the D8 gradient, which is on the path;
  * closed-form Fourier symbols of D8 and L3 and the exact growth factor of the
    linearised passive step;
  * a literal loop restatement of the Tjhung CPU scheme (test/tjhung_ambplus.h:103-251)
    and the reference's cross-implementation check (test/test_tjhung.cc:8-41);
  * an independent NumPy restatement (tests/numpy_scheme.py) on tiny grids;
  * invariants: mass, constant fields, shift and transposition equivariance.
"""
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import numpy_scheme as nps

D8 = {1: 4 / 5, 2: -1 / 5, 3: 4 / 105, 4: -1 / 280}


# ---- Philox ------------------------------------------------------------------

RANDOM123_KAT = [
    # Random123 kat_vectors, philox4x32 R=10: counter, key -> output
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expected", RANDOM123_KAT)
def test_philox_known_answers(oracle, ctr, key, expected):
    assert oracle.philox4x32_10(ctr, key) == expected


ROCRAND_PROBE = r"""
#include <cstdio>
#include <cstdlib>
#include <rocrand/rocrand_philox4x32_10.h>
struct Probe : rocrand_device::philox4x32_10_engine {
    Probe() : rocrand_device::philox4x32_10_engine(0, 0, 0) {}
    uint4 rounds(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};
int main(int argc, char** argv) {
    Probe p;
    for (int i = 1; i + 5 < argc + 1; i += 6) {
        uint4 c; uint2 k;
        c.x = strtoul(argv[i], 0, 16); c.y = strtoul(argv[i+1], 0, 16);
        c.z = strtoul(argv[i+2], 0, 16); c.w = strtoul(argv[i+3], 0, 16);
        k.x = strtoul(argv[i+4], 0, 16); k.y = strtoul(argv[i+5], 0, 16);
        uint4 o = p.rounds(c, k);
        printf("%08x %08x %08x %08x\n", o.x, o.y, o.z, o.w);
    }
    return 0;
}
"""


def test_philox_matches_rocrand(oracle):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc) or not os.path.exists("/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"):
        pytest.skip("hipcc / rocRAND headers not available")
    rng = np.random.default_rng(7)
    cases = [tuple(int(x) for x in rng.integers(0, 2**32, size=6, dtype=np.uint64)) for _ in range(16)]
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "probe.cpp"), os.path.join(d, "probe")
        open(src, "w").write(ROCRAND_PROBE)
        r = subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", src, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("could not build the rocRAND probe: " + r.stderr[-300:])
        args = [f"{x:08x}" for case in cases for x in case]
        out = subprocess.run([exe, *args], capture_output=True, text=True, check=True).stdout.split("\n")
    for case, line in zip(cases, out):
        got = oracle.philox4x32_10(case[:4], case[4:])
        assert tuple(int(w, 16) for w in line.split()) == got


def test_noise_mapping_and_moments(oracle):
    ny, nx, seed, step = 24, 36, 12345, 77
    xi = oracle.noise(ny, nx, seed, step)
    # group mapping: column j of row i, component c comes from group j//4, lane j%4
    for (c, i, j) in [(0, 0, 0), (1, 5, 7), (0, 23, 35), (1, 11, 18)]:
        z = oracle.noise_group(seed, step, c, i, j // 4)
        assert xi[c, i, j] == z[j % 4]
    # Box-Muller by hand for one group
    ctr = (3, 2, step & 0xFFFFFFFF, ((step >> 32) & 0x7FFFFFFF) | (1 << 31))
    w = oracle.philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))
    u1, u2 = (w[0] + 0.5) * 2.0**-32, w[1] * 2.0**-32
    r = math.sqrt(-2 * math.log(u1))
    assert xi[1, 2, 12] == pytest.approx(r * math.cos(2 * math.pi * u2), abs=1e-15)
    assert xi[1, 2, 13] == pytest.approx(r * math.sin(2 * math.pi * u2), abs=1e-15)
    # moments over many sites / steps, and independence of components and steps
    big = np.concatenate([oracle.noise(128, 128, 99, s).ravel() for s in range(8)])
    assert abs(big.mean()) < 5 / math.sqrt(big.size)
    assert abs(big.var() - 1) < 6 * math.sqrt(2 / big.size)
    assert abs(np.mean(big**4) - 3) < 0.05
    a, b = oracle.noise(64, 64, 99, 0), oracle.noise(64, 64, 99, 1)
    assert abs(np.corrcoef(a[0].ravel(), a[1].ravel())[0, 1]) < 0.03
    assert abs(np.corrcoef(a[0].ravel(), b[0].ravel())[0, 1]) < 0.03


# ---- the reference's polynomial KATs (D8 gradient, on the path) ----------------

@pytest.mark.parametrize("kind", ["x", "y", "xx", "yy"])
def test_d8_gradient_polynomial_kat(oracle, kind):
    """test/test_finite_differences.cc:48-121 at the centre of a 32x64 grid."""
    ny, nx = 32, 64
    i, j = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    phi = {"x": j, "y": i, "xx": 0.5 * j**2, "yy": 0.5 * i**2}[kind].astype(float)
    # mu_p = phi with a = 1, so J_p = -(D8y phi / dy, D8x phi / dx)
    t = oracle.terms(phi, (1e-2, 1, 1), (1, 0, 0, 0, 0, 0, 0))
    gy, gx = -t["jp"][0][ny // 2, nx // 2], -t["jp"][1][ny // 2, nx // 2]
    expected = {"x": (0, 1), "y": (1, 0), "xx": (0, nx / 2), "yy": (ny / 2, 0)}[kind]
    assert gy == pytest.approx(expected[0], abs=1e-12)
    assert gx == pytest.approx(expected[1], abs=1e-12)


# ---- closed forms ------------------------------------------------------------------

def d8_symbol(k):
    return 2 * sum(w * math.sin(m * k) for m, w in D8.items())


def l3_symbol(kx, ky, dx, dy):
    return ((1.5 - 0.5 * math.cos(ky)) * (2 * math.cos(kx) - 2) / dx**2
            + (1.5 - 0.5 * math.cos(kx)) * (2 * math.cos(ky) - 2) / dy**2)


@pytest.mark.parametrize("mx,my,dx,dy", [(1, 0, 1, 1), (3, 5, 1, 1), (7, 2, 1, 0.75), (16, 8, 1.3, 1.25)])
def test_fourier_symbols(oracle, mx, my, dx, dy):
    ny, nx = 32, 48
    kx, ky = 2 * math.pi * mx / nx, 2 * math.pi * my / ny
    i, j = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    th = kx * j + ky * i
    phi = np.cos(th)
    # L3 through mu_p = -kappa L3 phi
    t = oracle.terms(phi, (1e-2, dx, dy), (0, 0, 0, 1, 0, 0, 0))
    np.testing.assert_allclose(t["mup"], -l3_symbol(kx, ky, dx, dy) * phi, atol=1e-12)
    # D8 through J_p = -D8 mu_p with mu_p = phi
    t = oracle.terms(phi, (1e-2, dx, dy), (1, 0, 0, 0, 0, 0, 0))
    np.testing.assert_allclose(t["jp"][1], d8_symbol(kx) / dx * np.sin(th), atol=1e-12)
    np.testing.assert_allclose(t["jp"][0], d8_symbol(ky) / dy * np.sin(th), atol=1e-12)
    # the symbol peaks at ~1.7306 near k = 2.03 and vanishes at 0 and pi
    assert max(d8_symbol(k) for k in np.linspace(0, math.pi, 2001)) == pytest.approx(1.7306, abs=1e-3)
    assert abs(d8_symbol(math.pi)) < 1e-15


@pytest.mark.parametrize("mx,my,nsteps", [(2, 1, 1), (5, 3, 7), (11, 0, 3)])
def test_linear_growth_factor(oracle, mx, my, nsteps):
    """b=c=lambda=zeta=T=0: phi-hat <- (1 - dt Gamma (a - kappa L3-hat)) phi-hat exactly."""
    ny, nx, dt, dx, dy, a, kappa, eps, phi0 = 24, 40, 1e-2, 1.0, 0.8, -0.25, 1.0, 1e-3, 0.3
    kx, ky = 2 * math.pi * mx / nx, 2 * math.pi * my / ny
    i, j = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    mode = np.cos(kx * j + ky * i)
    phi = phi0 + eps * mode
    out = oracle.run(phi, (dt, dx, dy), (a, 0, 0, kappa, 0, 0, 0), nsteps)
    gamma = d8_symbol(kx) ** 2 / dx**2 + d8_symbol(ky) ** 2 / dy**2
    g = 1 - dt * gamma * (a - kappa * l3_symbol(kx, ky, dx, dy))
    np.testing.assert_allclose(out, phi0 + eps * g**nsteps * mode, atol=1e-14)


# ---- the reference's own CPU scheme, restated literally ----------------------------

def tjhung_calculate_dphi(phi, dt, dx, dy, A, K, lam, zeta):
    """Loop restatement of test/tjhung_ambplus.h:103-251 (T = 0).  Its first index
    (its 'x') is our row index; returns mueq, muact, Jeq, Jact and dphi."""
    Nx, Ny = phi.shape
    up = lambda i, n: (i + 1) % n
    dn = lambda i, n: (i - 1) % n

    def diff_x(arr, i, j):
        iu, idn, ju, jd = up(i, Nx), dn(i, Nx), up(j, Ny), dn(j, Ny)
        return ((arr[iu, ju] - arr[idn, ju]) * 0.1 + (arr[iu, j] - arr[idn, j]) * 0.3
                + (arr[iu, jd] - arr[idn, jd]) * 0.1) / dx

    def diff_y(arr, i, j):
        iu, idn, ju, jd = up(i, Nx), dn(i, Nx), up(j, Ny), dn(j, Ny)
        return ((arr[iu, ju] - arr[iu, jd]) * 0.1 + (arr[i, ju] - arr[i, jd]) * 0.3
                + (arr[idn, ju] - arr[idn, jd]) * 0.1) / dy

    def laplacian(arr, i, j):
        iu, idn, ju, jd = up(i, Nx), dn(i, Nx), up(j, Ny), dn(j, Ny)
        return ((-0.5 * arr[idn, ju] + 2.0 * arr[i, ju] - 0.5 * arr[iu, ju])
                + (2.0 * arr[idn, j] - 6.0 * arr[i, j] + 2.0 * arr[iu, j])
                + (-0.5 * arr[idn, jd] + 2.0 * arr[i, jd] - 0.5 * arr[iu, jd])) / (dx * dy)

    def diff_x_8(arr, i, j):
        s = sum(w * (arr[(i + m) % Nx, j] - arr[(i - m) % Nx, j]) for m, w in D8.items())
        return s / dx

    def diff_y_8(arr, i, j):
        s = sum(w * (arr[i, (j + m) % Ny] - arr[i, (j - m) % Ny]) for m, w in D8.items())
        return s / dy

    shape = phi.shape
    dpx, dpy, lap, mueq, muact = (np.zeros(shape) for _ in range(5))
    for i in range(Nx):
        for j in range(Ny):
            dpx[i, j], dpy[i, j] = diff_x(phi, i, j), diff_y(phi, i, j)
    for i in range(Nx):
        for j in range(Ny):
            lap[i, j] = laplacian(phi, i, j)
            mueq[i, j] = -A * phi[i, j] + A * phi[i, j] ** 3 - K * lap[i, j]
            muact[i, j] = lam * (dpx[i, j] ** 2 + dpy[i, j] ** 2)
    jxeq, jyeq, jxact, jyact = (np.zeros(shape) for _ in range(4))
    for i in range(Nx):
        for j in range(Ny):
            jxeq[i, j], jyeq[i, j] = -diff_x_8(mueq, i, j), -diff_y_8(mueq, i, j)
            jxact[i, j] = zeta * lap[i, j] * dpx[i, j] - diff_x(muact, i, j)
            jyact[i, j] = zeta * lap[i, j] * dpy[i, j] - diff_y(muact, i, j)
    dphi = np.zeros(shape)
    for i in range(Nx):
        for j in range(Ny):
            dphi[i, j] = -(diff_x_8(jxeq, i, j) + diff_y_8(jyeq, i, j)
                           + diff_x(jxact, i, j) + diff_y(jyact, i, j))
    return dict(mueq=mueq, muact=muact, jeq=(jxeq, jyeq), jact=(jxact, jyact), dphi=dphi)


def test_oracle_matches_tjhung_scheme(oracle):
    """test/test_tjhung.cc:8-41: 64x32 in the reference, A=0.25, kappa=lambda=zeta=1,
    T=0, dt=1e-2; here 16x12 to keep the Python loops short."""
    rng = np.random.default_rng(3)
    phi = rng.uniform(-1, 1, size=(16, 12))
    A, dt = 0.25, 1e-2
    ref = tjhung_calculate_dphi(phi, dt, 1.0, 1.0, A, 1.0, 1.0, 1.0)
    t = oracle.terms(phi, (dt, 1, 1), (-A, 0, A, 1, 1, 1, 0))
    np.testing.assert_allclose(t["mup"], ref["mueq"], atol=1e-12, rtol=0)
    np.testing.assert_allclose(t["mua"], ref["muact"], atol=1e-12, rtol=0)
    np.testing.assert_allclose(t["jp"][0], ref["jeq"][0], atol=1e-12, rtol=0)
    np.testing.assert_allclose(t["jp"][1], ref["jeq"][1], atol=1e-12, rtol=0)
    np.testing.assert_allclose(t["jl"][0] + t["jz"][0], ref["jact"][0], atol=1e-12, rtol=0)
    np.testing.assert_allclose(t["jl"][1] + t["jz"][1], ref["jact"][1], atol=1e-12, rtol=0)
    np.testing.assert_allclose(-t["divj"], ref["dphi"], atol=1e-12, rtol=0)
    out = oracle.run(phi, (dt, 1, 1), (-A, 0, A, 1, 1, 1, 0), 1)
    np.testing.assert_allclose(out, phi + dt * ref["dphi"], atol=1e-14, rtol=0)


# ---- independent NumPy restatement ------------------------------------------------

@pytest.mark.parametrize("shape", [(8, 8), (12, 20), (17, 9), (5, 3), (1, 7)])
@pytest.mark.parametrize("stencil,model", [
    ((1e-2, 1, 1), (-0.25, 0.5, 0.25, 1, 1, 1, 0)),
    ((1e-2, 1, 0.75), (1, 2, 3, 0.5, -1, -4, 0.1)),
    ((5e-3, 1.25, 1), (-0.25, 0, 0.25, 1, 0.5, -1, 0.05)),
])
def test_oracle_matches_numpy_restatement(oracle, shape, stencil, model):
    rng = np.random.default_rng(hash((shape, stencil)) % 2**32)
    phi = rng.normal(0, 0.3, size=shape)
    xi = rng.normal(size=(2,) + shape)
    t = oracle.terms(phi, stencil, model, xi=xi)
    ref = nps.terms(phi, stencil, model, xi=xi)
    for k in ("mup", "mua", "jp", "jl", "jz", "jn", "divj"):
        np.testing.assert_allclose(t[k], ref[k], atol=1e-12, rtol=1e-12, err_msg=k)
    out = oracle.run(phi, stencil, model, 1, noise_buf=xi[None])
    np.testing.assert_allclose(out, nps.step(phi, stencil, model, xi), atol=1e-13)


def test_oracle_run_uses_philox_noise(oracle):
    rng = np.random.default_rng(11)
    phi = rng.normal(0, 0.1, size=(12, 16))
    st, md, seed = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1), 42
    buf = np.stack([oracle.noise(12, 16, seed, 10 + s) for s in range(3)])
    a = oracle.run(phi, st, md, 3, seed=seed, step0=10)
    b = oracle.run(phi, st, md, 3, noise_buf=buf)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, oracle.run(phi, st, md, 3, seed=seed + 1, step0=10))


# ---- invariants ---------------------------------------------------------------------

def test_invariants(oracle):
    rng = np.random.default_rng(5)
    ny, nx = 20, 28
    phi = rng.normal(0, 0.3, size=(ny, nx))
    st, md = (1e-2, 1, 1), (-0.25, 0.5, 0.25, 1, 1.5, -2, 0)
    t = oracle.terms(phi, st, md)
    assert abs(t["divj"].sum()) < 1e-11 * np.abs(t["divj"]).sum()
    # constant field is a fixed point at T = 0
    assert np.array_equal(oracle.run(np.full((ny, nx), 0.37), st, md, 5), np.full((ny, nx), 0.37))
    # shift equivariance
    s = np.roll(np.roll(phi, 3, 0), -5, 1)
    np.testing.assert_allclose(oracle.terms(s, st, md)["divj"],
                               np.roll(np.roll(t["divj"], 3, 0), -5, 1), atol=1e-13)
    # transposition (square spacing): x and y exchange roles
    sq = phi[:, :ny]
    a = oracle.terms(sq, st, md)["divj"]
    b = oracle.terms(np.ascontiguousarray(sq.T), st, md)["divj"]
    np.testing.assert_allclose(b, a.T, atol=1e-12)
    # mass over many steps with noise
    out = oracle.run(phi, (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1), 50, seed=1)
    assert abs(out.sum() - phi.sum()) < 1e-11


def test_multithreaded_oracle_is_identical(oracle):
    rng = np.random.default_rng(9)
    phi = rng.normal(0, 0.1, size=(64, 48))
    st, md = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1)
    a = oracle.run(phi, st, md, 5, seed=3, nthreads=1)
    b = oracle.run(phi, st, md, 5, seed=3, nthreads=4)
    assert np.array_equal(a, b)
