"""Saving and loading the autotuners' choices (no reference counterpart).

``Integrator.autotune`` and ``Ensemble.autotune`` time every compiled tile shape and
record the fastest in a process-wide table keyed by (precision, ny, nx, noisy); new
integrators and ensembles of a tabled size start from its entry instead of the
occupancy rule, and ``autotune(cached=True)`` returns it without timing.  This module
keeps that table across processes as JSON.  Entries are stored by tile shape
(columns, rows, threads), not by index, so a file written by one build loads into
another; shapes this build does not compile are skipped.

    from activemodelbplus import tuning
    tuning.load("tiles.json")        # before creating integrators
    sim.autotune(4, cached=True)     # times only sizes the file does not cover
    tuning.save("tiles.json")
"""
from __future__ import annotations

import json
import os

from .integrator import tile_shapes
from .integrator import tuning_clear as clear  # noqa: F401  (re-exported)
from .integrator import tuning_entries, tuning_set

FORMAT = 1


def entries():
    """The table as a list of dicts: precision, ny, nx, noisy, tile (cols, rows, threads)."""
    out = []
    for precision, ny, nx, noisy, tile in tuning_entries():
        out.append({"precision": int(precision), "ny": int(ny), "nx": int(nx), "noisy": bool(noisy),
                    "tile": list(tile_shapes(precision)[tile])})
    return out


def save(path):
    """Writes the table to `path` (JSON); returns the number of entries written."""
    data = {"format": FORMAT, "entries": entries()}
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, path)
    return len(data["entries"])


def load(path):
    """Adds the entries of a file written by `save` to the table (existing entries of the
    same key are replaced); returns the number loaded.  A missing file loads nothing."""
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        data = json.load(f)
    if data.get("format") != FORMAT:
        raise ValueError(f"{path}: unknown tuning file format {data.get('format')!r}")
    shapes = {p: [list(s) for s in tile_shapes(p)] for p in (32, 64)}
    n = 0
    for e in data.get("entries", []):
        p = int(e["precision"])
        if p not in shapes or list(e["tile"]) not in shapes[p]:
            continue  # a shape this build does not compile
        tuning_set(p, int(e["ny"]), int(e["nx"]), bool(e["noisy"]), shapes[p].index(list(e["tile"])))
        n += 1
    return n
