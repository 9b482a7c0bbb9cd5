"""The autotuners' table and its JSON persistence (host-side; no GPU needed), and on the
GPU: autotune records its choice, new integrators and ensembles start from the table, and
autotune(cached=True) returns the entry without timing."""
import json

import numpy as np
import pytest


@pytest.fixture()
def tuning():
    from activemodelbplus import tuning as T
    T.clear()
    yield T
    T.clear()


def test_table_round_trip(tuning, tmp_path):
    import activemodelbplus.integrator as amb
    amb.tuning_set(32, 4096, 4096, True, 0)
    amb.tuning_set(32, 4096, 4096, False, 1)
    amb.tuning_set(64, 256, 256, False, 5)
    assert amb.tuning_get(32, 4096, 4096, True) == 0
    assert amb.tuning_get(32, 100, 100, True) == -1
    path = str(tmp_path / "tiles.json")
    assert tuning.save(path) == 3
    data = json.load(open(path))
    assert data["format"] == 1
    assert {(e["precision"], e["ny"], e["noisy"], tuple(e["tile"])) for e in data["entries"]} == {
        (32, 4096, True, tuple(amb.tile_shapes(32)[0])), (32, 4096, False, tuple(amb.tile_shapes(32)[1])),
        (64, 256, False, tuple(amb.tile_shapes(64)[5]))}
    tuning.clear()
    assert amb.tuning_entries() == []
    # a shape this build does not compile is skipped
    data["entries"].append({"precision": 32, "ny": 8, "nx": 8, "noisy": True, "tile": [24, 24, 100]})
    json.dump(data, open(path, "w"))
    assert tuning.load(path) == 3
    assert amb.tuning_get(64, 256, 256, False) == 5
    assert tuning.load(str(tmp_path / "missing.json")) == 0


def test_bad_entries_rejected(tuning):
    import activemodelbplus.integrator as amb
    with pytest.raises(ValueError):
        amb.tuning_set(16, 8, 8, True, 0)
    with pytest.raises(ValueError):
        amb.tuning_set(32, 8, 8, True, 99)


@pytest.mark.gpu
def test_autotune_records_and_new_handles_start_from_it(tuning):
    import activemodelbplus.integrator as amb
    rng = np.random.default_rng(3)
    phi = 0.05 * rng.normal(size=(320, 288))
    st, md = amb.Stencil(5e-3, 1, 1), amb.Model(-0.25, 0, 0.25, 1, 0.5, -1, 0.05)
    sim = amb.Integrator(phi, st, md, seed=1, precision=32)
    best = sim.autotune(2)
    assert amb.tuning_get(32, 320, 288, True) == best
    # a different tile in the table is taken by new integrators and by cached autotune
    other = (best + 1) % len(amb.tile_shapes(32))
    amb.tuning_set(32, 320, 288, True, other)
    fresh = amb.Integrator(phi, st, md, seed=1, precision=32)
    assert fresh.options["tile"] - 1 == other
    assert fresh.autotune(2, cached=True) == other
    ens = amb.Ensemble([phi, phi[:64, :64]], st, md, [1, 2], precision=32)
    assert ens.tiles[0] == other
    # results do not depend on the tile
    sim.run(40)
    fresh.run(40)
    assert np.array_equal(sim.field, fresh.field)
