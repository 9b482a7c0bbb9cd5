"""GPU parity: the HIP path (through the C ABI and the pybind11 module) against the
CPU oracle on the same inputs.  Mirrors the reference's GPU tests
(test/test_integrator.cc, test/test_tjhung.cc) and adds the multi-step parity,
noise, precision, mass and edge-case checks.

Bars: FP64 per-term getters and single steps to 1e-12 absolute (the reference's
tight_tol, test/test_utilities.h:7); FP64 trajectories to 1e-10 relative max-norm;
FP32 against FP64 within the bounds stated per test.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TIGHT = 1e-12


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def amb():
    import activemodelbplus.integrator as m
    return m


def make(amb, phi, stencil, model, **kw):
    return amb.Integrator(phi, amb.Stencil(*stencil), amb.Model(*model), **kw)


# ---- the reference's per-term and cross-implementation tests --------------------

def test_tjhung_cross_implementation(amb, oracle):
    """test/test_tjhung.cc:8-41: 64x32, dt=1e-2, A=0.25, kappa=lambda=zeta=1, T=0."""
    rng = np.random.default_rng(0)
    phi = rng.uniform(-1, 1, size=(64, 32))
    st, md = (1e-2, 1, 1), (-0.25, 0, 0.25, 1, 1, 1, 0)
    sim = make(amb, phi, st, md, seed=1)
    ref = oracle.terms(phi, st, md)
    np.testing.assert_allclose(sim.passive_chemical_potential, ref["mup"], atol=TIGHT, rtol=0)
    np.testing.assert_allclose(sim.active_chemical_potential, ref["mua"], atol=TIGHT, rtol=0)
    for c in range(2):
        np.testing.assert_allclose(sim.passive_current[c], ref["jp"][c], atol=TIGHT, rtol=0)
        np.testing.assert_allclose(sim.active_current[c], ref["jl"][c] + ref["jz"][c], atol=TIGHT, rtol=0)
    sim.run(1)
    np.testing.assert_allclose(-(sim.field - phi) / st[0], ref["divj"], atol=TIGHT, rtol=0)


@pytest.mark.parametrize("name,stencil,model", [
    ("bulk", (1e-2, 1, 0.75), (1, 2, 3, 0, 0, 0, 0)),          # test_integrator.cc:47-80
    ("kappa", (1e-2, 1, 0.75), (0, 0, 0, 1, 0, 0, 0)),         # :82-111
    ("lambda", (1e-2, 1, 0.75), (0, 0, 0, 0, 1, 0, 0)),        # :113-149
    ("zeta", (1e-2, 1, 1.25), (0, 0, 0, 0, 0, 1, 0)),          # :151-195
])
def test_per_term_currents(amb, oracle, name, stencil, model):
    rng = np.random.default_rng(1)
    phi = rng.uniform(-1, 1, size=(32, 64))
    sim = make(amb, phi, stencil, model, seed=2)
    ref = oracle.terms(phi, stencil, model)
    np.testing.assert_allclose(sim.chemical_potential, ref["mup"] + ref["mua"], atol=TIGHT, rtol=0)
    for c in range(2):
        np.testing.assert_allclose(sim.passive_current[c], ref["jp"][c], atol=TIGHT, rtol=0)
        np.testing.assert_allclose(sim.lambda_current[c], ref["jl"][c], atol=TIGHT, rtol=0)
        np.testing.assert_allclose(sim.circulating_current[c], ref["jz"][c], atol=TIGHT, rtol=0)
        np.testing.assert_allclose(sim.random_current[c], 0.0, atol=0)
    sim.run(1)
    np.testing.assert_allclose(-(sim.field - phi) / stencil[0], ref["divj"], atol=TIGHT, rtol=0)


# ---- multi-step parity (1e-10 relative max-norm) -----------------

def spinodal(n, seed):
    rng = np.random.default_rng(seed)
    phi = 0.1 * rng.random((n, n))
    return phi - phi.mean()


def test_config1_fp64_parity_1000_steps(amb, oracle):
    """Passive spinodal, 256^2, FP64, dt=1e-3: 10^3 steps vs the oracle."""
    phi = spinodal(256, 1)
    st, md = (1e-3, 1, 1), (-0.25, 0, 0.25, 1, 0, 0, 0)
    sim = make(amb, phi, st, md, seed=1)
    sim.run(1000)
    ref = oracle.run(phi, st, md, 1000, nthreads=8)
    err = relmax(sim.field, ref)
    print(f"config1 FP64 1e3 steps: rel max-norm error {err:.3e}")
    assert err <= 1e-10
    s, bad = sim.mass()
    assert bad == 0 and abs(s - phi.sum()) < 1e-9


def test_config1_fp64_parity_full_length(amb, oracle):
    """Config 1 at its full length: passive spinodal, 256^2, FP64, dt=1e-3, 10^4 steps
    vs the oracle (the CPU scheme on the host's cores)."""
    phi = spinodal(256, 1)
    st, md = (1e-3, 1, 1), (-0.25, 0, 0.25, 1, 0, 0, 0)
    sim = make(amb, phi, st, md, seed=1)
    sim.run(10_000)
    ref = oracle.run(phi, st, md, 10_000, nthreads=min(16, oracle.max_threads()))
    err = relmax(sim.field, ref)
    print(f"config1 FP64 1e4 steps: rel max-norm error {err:.3e}")
    assert err <= 1e-10
    s, bad = sim.mass()
    assert bad == 0 and abs(s - phi.sum()) < 1e-9


def test_config2_fp32_against_fp64_full_length(amb):
    """Config 2 at its full length (1024^2, lamb=-1, zeta=-4, T=0.1, dt=5e-3, 10^5 steps,
    phi = -0.4 + 0.01 N(0,1)): the FP32 path against the FP64 path on the same Philox
    stream.  The noisy trajectories decorrelate pointwise long before 10^5 steps, so the
    comparison is statistical: mass, finiteness and the structure factor's first moment."""
    from activemodelbplus import trajectory as T
    rng = np.random.default_rng(2)
    phi = -0.4 + 0.01 * rng.normal(size=(1024, 1024))
    st, md = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1)
    out = {}
    for prec in (64, 32):
        sim = make(amb, phi, st, md, seed=2, precision=prec)
        sim.run(100_000)
        s, bad = sim.mass()
        out[prec] = (sim, s, bad)
        assert bad == 0
    k64, k32 = T.mean_wavenumber(out[64][0]), T.mean_wavenumber(out[32][0])
    f64, f32 = out[64][0].field, out[32][0].field
    l2 = float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64))
    m0 = float(np.sum(phi))
    print(f"config2 1e5 steps: <k> FP64 {k64:.4f} FP32 {k32:.4f} ({k32 / k64 - 1:+.3%}); "
          f"L2(FP32 - FP64)/L2(FP64) {l2:.3e}; mass drift FP64 {out[64][1] - m0:.3e} FP32 {out[32][1] - m0:.3e}")
    assert abs(out[64][1] - m0) < 1e-6
    assert abs(out[32][1] - m0) / phi.size < 1e-5
    assert abs(k32 / k64 - 1) < 0.05


@pytest.mark.parametrize("shape,stencil,model", [
    ((64, 96), (1e-2, 1, 1), (-0.25, 0.5, 0.25, 1, 1, 1, 0)),
    ((48, 40), (5e-3, 1, 1.25), (-0.25, 0, 0.25, 1, -1, -4, 0)),
    ((37, 53), (1e-2, 1, 0.75), (-0.25, 0.5, 0.25, 1, 1, 1, 0)),    # not tile multiples
    ((7, 5), (1e-2, 1, 1), (-0.25, 0.5, 0.25, 1, 0.5, -1, 0)),        # smaller than a halo
    ((1, 16), (1e-2, 1, 1), (-0.25, 0, 0.25, 1, 1, 1, 0)),
])
def test_active_fp64_parity(amb, oracle, shape, stencil, model):
    rng = np.random.default_rng(4)
    phi = -0.2 + 0.1 * rng.normal(size=shape)
    sim = make(amb, phi, stencil, model, seed=3)
    sim.run(1000)
    ref = oracle.run(phi, stencil, model, 1000)
    assert relmax(sim.field, ref) <= 1e-10


@pytest.mark.parametrize("shape", [(64, 64), (40, 36), (33, 30)])
def test_noise_shared_buffer_parity(amb, oracle, shape):
    """T > 0 with both sides reading one shared buffer of normals: 1e-10 relative."""
    rng = np.random.default_rng(5)
    phi = -0.4 + 0.01 * rng.normal(size=shape)
    st, md = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1)
    nsteps = 400
    buf = rng.normal(size=(nsteps, 2) + shape)
    sim = make(amb, phi, st, md, seed=9)
    sim.set_noise(buf)
    np.testing.assert_allclose(np.stack(sim.random_current),
                               math.sqrt(2 * 0.1 / 5e-3) * buf[0], atol=1e-13, rtol=0)
    sim.run(nsteps)
    ref = oracle.run(phi, st, md, nsteps, noise_buf=buf)
    assert relmax(sim.field, ref) <= 1e-10


@pytest.mark.parametrize("shape", [(64, 128), (30, 44), (21, 19)])
def test_philox_noise_parity(amb, oracle, shape):
    """The device Philox stream equals the oracle's (nx % 4 != 0 exercises the seam)."""
    rng = np.random.default_rng(6)
    phi = -0.4 + 0.01 * rng.normal(size=shape)
    st, md, seed = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1), 2024
    sim = make(amb, phi, st, md, seed=seed, timestep=17)
    amp = math.sqrt(2 * 0.1) * math.sqrt(1 / 5e-3)
    np.testing.assert_allclose(np.stack(sim.random_current), amp * oracle.noise(*shape, seed, 17),
                               atol=1e-12, rtol=0)
    sim.run(200)
    ref = oracle.run(phi, st, md, 200, seed=seed, step0=17)
    assert relmax(sim.field, ref) <= 1e-10
    assert sim.timestep == 217


def _random_case(i):
    """Case i of the randomized parity sweep (seeded): shape 1..80 per side, spacings,
    model parameters in the ranges the reference's tests and examples use, T >= 0."""
    rng = np.random.default_rng(1000 + i)
    ny, nx = (int(v) for v in rng.integers(1, 81, size=2))
    dx, dy = (float(v) for v in rng.choice([0.75, 1.0, 1.25], size=2))
    dt = float(rng.choice([1e-3, 2e-3]))
    model = (-0.25, float(rng.uniform(-0.5, 0.5)), 0.25, float(rng.uniform(0.5, 1.5)), float(rng.uniform(-2, 2)),
             float(rng.uniform(-5, 1)), float(rng.choice([0.0, 0.05, 0.1])))
    phi = float(rng.uniform(-0.5, 0.5)) + 0.05 * rng.normal(size=(ny, nx))
    return phi, (dt, dx, dy), model, int(rng.integers(1, 2**40))


@pytest.mark.parametrize("i", range(24))
def test_randomized_fp64_parity(amb, oracle, i):
    """Seeded random shapes (tiny, odd, seam-wrapping), spacings, parameters and
    temperatures: 40 FP64 steps with the Philox stream on both sides vs the oracle."""
    phi, st, md, seed = _random_case(i)
    sim = make(amb, phi, st, md, seed=seed, timestep=i)
    sim.run(40)
    ref = oracle.run(phi, st, md, 40, seed=seed, step0=i)
    err = relmax(sim.field, ref)
    assert err <= 1e-10, (phi.shape, st, md, err)


@pytest.mark.parametrize("precision", [32, 64])
def test_randomized_ensemble_members_equal_single_runs(amb, precision):
    """12 members of random sizes (8..100 per side, some not multiples of 4), parameters,
    temperatures and seeds advanced together equal plain integrators bit for bit."""
    rng = np.random.default_rng(77 + precision)
    fields, models, seeds = [], [], []
    for _ in range(12):
        ny, nx = (int(v) for v in rng.integers(8, 101, size=2))
        fields.append(float(rng.uniform(-0.5, 0.5)) + 0.05 * rng.normal(size=(ny, nx)))
        models.append(amb.Model(-0.25, 0, 0.25, 1, float(rng.uniform(-2, 2)), float(rng.uniform(-5, 0)),
                                float(rng.choice([0.0, 0.05, 0.1]))))
        seeds.append(int(rng.integers(1, 2**40)))
    st = amb.Stencil(2e-3, 1, 1)
    ens = amb.Ensemble(fields, st, models, seeds, precision=precision)
    ens.run(30)
    for i in range(12):
        sim = amb.Integrator(fields[i], st, models[i], seed=seeds[i], precision=precision)
        sim.run(30)
        np.testing.assert_array_equal(ens.field(i), sim.field, err_msg=f"member {i} {fields[i].shape}")


def test_fp32_against_fp64(amb, oracle):
    """FP32 state and arithmetic vs the FP64 oracle (noisy active config, 128^2)."""
    rng = np.random.default_rng(2)
    phi = -0.4 + 0.01 * rng.normal(size=(128, 128))
    st, md, seed = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1), 2
    sim = make(amb, phi, st, md, seed=seed, precision=32)
    sim.run(500)
    ref = oracle.run(phi, st, md, 500, seed=seed, nthreads=8)
    err = relmax(sim.field, ref)
    print(f"FP32 vs FP64 oracle, 128^2, 500 noisy steps: rel max-norm {err:.3e}")
    assert err < 1e-4
    s, bad = sim.mass()
    assert bad == 0 and abs(s - phi.sum()) / phi.size < 1e-6


@pytest.mark.parametrize("temperature", [0.1, 0.0])
def test_tile_shapes_bit_identical(amb, temperature):
    """Every compiled tile shape computes each site with the same arithmetic, in the noisy
    and in the noiseless kernel."""
    rng = np.random.default_rng(8)
    for precision in (64, 32):
        phi = -0.3 + 0.05 * rng.normal(size=(96, 160))
        out = []
        for t in range(len(amb.tile_shapes(precision))):
            sim = make(amb, phi, (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, temperature), seed=5,
                       precision=precision, tile=t + 1)
            sim.run(40)
            out.append(sim.field)
        for o in out[1:]:
            assert np.array_equal(o, out[0])


def test_active_current_fields_tiles_members_and_fp64(amb):
    """FP32 noisy steps on fields of 2048^2 sites or more form the G3 current of the active
    terms in the noise stage (smaller fields and FP64 apply the 5x5 stencil of mu_a in the
    last stage): at 2048^2 every tile shape still gives the same bits, a batched member equals
    the plain integrator, and the field stays within FP32 rounding of the FP64 path."""
    st, md = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, 0.5, -1, 0.05)
    phi = 0.05 * np.random.default_rng(9).normal(size=(2048, 2048))
    out = []
    for t in range(len(amb.tile_shapes(32))):
        sim = make(amb, phi, st, md, seed=5, precision=32, tile=t + 1)
        sim.run(20)
        out.append(sim.field)
    for t, o in enumerate(out[1:], start=2):
        assert np.array_equal(o, out[0]), amb.tile_shapes(32)[t - 1]
    small = phi[:64, :96].copy()
    ens = amb.Ensemble([phi, small], amb.Stencil(*st), [amb.Model(*md)] * 2, [5, 6], precision=32)
    ens.run(20)
    np.testing.assert_array_equal(ens.field(0), out[0])
    ref = make(amb, phi, st, md, seed=5, precision=64)
    ref.run(20)
    err = relmax(out[0], ref.field)
    print(f"2048^2 noisy, 20 steps: FP32 (active current) vs FP64 rel max-norm {err:.3e}")
    assert err < 1e-5
    # the shared-buffer noise mode takes the same path
    buf = np.random.default_rng(10).normal(size=(3, 2) + phi.shape)
    runs = []
    for precision in (32, 64):
        sim = make(amb, phi, st, md, seed=5, precision=precision)
        sim.set_noise(buf)
        sim.run(3)
        runs.append(sim.field)
    err = relmax(runs[0], runs[1])
    print(f"2048^2 shared noise buffer, 3 steps: FP32 (active current) vs FP64 rel max-norm {err:.3e}")
    assert err < 1e-5
    # a field past the size whose columns do not split into groups of four (every tile takes
    # the per-site noise path)
    odd = 0.05 * np.random.default_rng(11).normal(size=(2050, 2050))
    shapes = amb.tile_shapes(32)
    runs = []
    for t, precision in ((1, 32), (len(shapes), 32), (0, 64)):
        sim = make(amb, odd, st, md, seed=7, precision=precision, tile=t)
        sim.run(10)
        runs.append(sim.field)
    assert np.array_equal(runs[0], runs[1])
    err = relmax(runs[0], runs[2])
    print(f"2050^2 noisy, 10 steps: FP32 (active current) vs FP64 rel max-norm {err:.3e}")
    assert err < 1e-5


def test_default_tile_per_noise_mode(amb):
    """Without autotuning the tile comes from the occupancy rule over each kernel's order of
    preference: at 4096^2 FP32 the noisy kernel takes 32x32 tiles of 512 threads (four
    workgroups per CU), the noiseless kernel 32x48 tiles of 512 threads; at 1024^2 one full
    round of 32x32 (noisy) or 64x32 tiles (noiseless), FP64 two rounds of 32x32 tiles of 512
    threads; fields too small for a round the shape that puts nearest 8 waves on each CU,
    the earlier in the order on a tie (32x32 tiles of 512 threads at 512^2 in both
    precisions, 16x8 of 128 threads at 256^2 and 128^2): the fastest shape measured there,
    or within 3% of it.  The autotuners' table (which earlier tests fill) is emptied for the
    check and restored after it."""
    from activemodelbplus import tuning
    saved = amb.tuning_entries()
    tuning.clear()
    try:
        _check_default_tiles(amb)
    finally:
        for e in saved:
            amb.tuning_set(*e)


def _check_default_tiles(amb):
    for precision, side, temperature, want in [(32, 4096, 0.05, (32, 32, 512)), (32, 4096, 0.0, (32, 32, 512)),
                                               (32, 1024, 0.1, (32, 32, 512)), (32, 1024, 0.0, (32, 32, 512)),
                                               (32, 512, 0.05, (32, 32, 512)), (32, 512, 0.0, (32, 32, 512)),
                                               (32, 256, 0.0, (16, 8, 128)), (32, 256, 0.05, (16, 8, 128)),
                                               (64, 256, 0.0, (16, 8, 128)), (64, 512, 0.05, (32, 32, 512)),
                                               (64, 512, 0.0, (32, 32, 512)), (64, 1024, 0.05, (32, 32, 512)),
                                               (32, 128, 0.0, (16, 8, 128)), (64, 128, 0.05, (16, 8, 128))]:
        sim = make(amb, np.zeros((side, side)), (5e-3, 1, 1), (-0.25, 0, 0.25, 1, 0.5, -1, temperature),
                   precision=precision)
        shape = tuple(amb.tile_shapes(precision)[sim.options["tile"] - 1])
        assert shape == want, (precision, side, temperature, shape)


def test_graph_and_direct_launch_agree(amb):
    """run(100) (graph replays + direct tail) equals 100 x run(1) (direct launches)."""
    rng = np.random.default_rng(10)
    phi = -0.3 + 0.05 * rng.normal(size=(64, 64))
    st, md = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1)
    a = make(amb, phi, st, md, seed=4)
    a.run(100)
    b = make(amb, phi, st, md, seed=4)
    for _ in range(100):
        b.run(1)
    assert np.array_equal(a.field, b.field)
    # replaying the same graph again keeps matching
    a.run(64)
    for _ in range(64):
        b.run(1)
    assert np.array_equal(a.field, b.field)


def test_checkpoint_resume_is_exact(amb, tmp_path):
    rng = np.random.default_rng(12)
    phi = -0.4 + 0.01 * rng.normal(size=(64, 64))
    st, md = (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1)
    a = make(amb, phi, st, md, seed=77)
    a.run(150)
    path = str(tmp_path / "ckpt.npz")
    a.save(path)
    a.run(150)
    b = amb.Integrator.from_checkpoint(path)
    assert b.timestep == 150
    b.run(150)
    assert np.array_equal(a.field, b.field)


# ---- bookkeeping, conservation, failure ------------------------------------------

def test_timestep_and_time(amb):
    """test/test_integrator.cc:253-266 and test/test_bindings.py:51-71."""
    sim = make(amb, np.random.default_rng(0).random((64, 64)), (0.1, 1, 1), (0, 0, 0, 0, 0, 0, 0))
    assert sim.timestep == 0
    sim.run(10)
    assert sim.timestep == 10
    assert sim.time == pytest.approx(1.0, abs=1e-12)
    st, md = amb.Stencil(1e-2, 1, 1), amb.Model(0.25, 0.5, 0.25, 1, 1, 1, 0)
    initial = np.random.random((128, 256))
    sim = amb.Integrator(initial, st, md)
    assert (sim.field == initial).all()
    assert sim.stencil is not st and sim.stencil == st and sim.model == md
    sim.run(1)
    assert (sim.field != initial).any()
    assert sim.timestep == 1 and sim.time == st.dt


def test_conservation_four_models(amb):
    """test/test_integrator.cc:197-228: 512x256, dt=1e-2, dy=1.25, 1000 steps."""
    rng = np.random.default_rng(13)
    initial = 0.1 * rng.random((256, 512))
    initial -= initial.mean()
    previous = None
    for md in [(-0.25, 0, 0.25, 1, 0, 0, 0), (-0.25, 0.5, 0.25, 1, 0, 0, 0),
               (-0.25, 0.5, 0.25, 1, 1, 0, 0), (-0.25, 0.5, 0.25, 1, 1, 1, 0)]:
        sim = make(amb, initial, (1e-2, 1, 1.25), md, seed=0)
        m0 = sim.field.sum()
        sim.run(1000)
        f = sim.field
        assert abs(f.sum() - m0) < 1e-6
        if previous is not None:
            assert np.max(np.abs(f - previous)) > 1e-6
        previous = f


def test_nonfinite_raises(amb):
    phi = np.zeros((32, 32))
    phi[3, 4] = np.nan
    sim = make(amb, phi, (1e-2, 1, 1), (-0.25, 0, 0.25, 1, 0, 0, 0))
    with pytest.raises(amb.CudaException):
        sim.run(1)
    # an unstable time step blows up and is reported the same way
    rng = np.random.default_rng(1)
    sim = make(amb, rng.normal(size=(32, 32)), (1.0, 1, 1), (-0.25, 0, 0.25, 1, 0, 0, 0))
    with pytest.raises(amb.CudaException):
        sim.run(200)


def test_run_for_time_and_interrupts(amb):
    sim = make(amb, spinodal(64, 3), (1e-2, 1, 1), (-0.25, 0, 0.25, 1, 0, 0, 0))
    seen = []
    sim.run_for_time(5.0, t_interrupt=[1.0, 2.5, 5.0], f_interrupt=lambda s: seen.append(s.time),
                     max_updates=4, min_update_time=0.5)
    assert sim.timestep == 500
    assert np.allclose(seen, [1.0, 2.5, 5.0])


def test_autotune_keeps_state(amb):
    phi = spinodal(256, 5)
    sim = make(amb, phi, (5e-3, 1, 1), (-0.25, 0, 0.25, 1, -1, -4, 0.1), seed=1, precision=32)
    before = sim.field
    best = sim.autotune(4)
    assert 0 <= best < len(amb.tile_shapes(32))
    assert np.array_equal(sim.field, before) and sim.timestep == 0


# ---- batched ensemble ---------------------------------------------------------------

def test_ensemble_matches_single_runs(amb):
    rng = np.random.default_rng(14)
    sides = [64, 128, 64, 96, 256, 50]
    fields, models, seeds = [], [], []
    for i, n in enumerate(sides):
        fields.append(rng.uniform(-0.6, 0.6) + 0.01 * rng.normal(size=(n, n)))
        models.append(amb.Model(-0.25, 0, 0.25, 1, rng.uniform(-2, 2), rng.uniform(-5, 0),
                                [0, 0.05, 0.1][i % 3]))
        seeds.append(1000 + i)
    st = amb.Stencil(5e-3, 1, 1)
    for precision in (32, 64):
        ens = amb.Ensemble(fields, st, models, seeds, precision=precision)
        ens.autotune(2)
        ens.run(70)
        for i in range(len(sides)):
            single = amb.Integrator(fields[i], st, models[i], seed=seeds[i], precision=precision)
            single.run(70)
            assert np.array_equal(ens.field(i), single.field), (precision, i)
            s, bad = ens.mass(i)
            assert bad == 0
