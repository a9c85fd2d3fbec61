"""Launch-shape tuner for the PCG of a staged solver.

The PCG's launch shapes (waves per CU of the FP64 and FP32 stencil kernels
and of the column kernels, include/frehg_gw.h frehg_gw_set_pcg_shapes) were
chosen by a sweep at 512 x 512 x 100; other grids, ensembles and devices may
prefer others.  ``tune(solver)`` times one step of the solver's own case per
candidate shape, every trial from the same state (head, theta and coupling
arrays are saved once and restored before each trial), by coordinate
descent over the three knobs, and leaves the fastest shapes set.  Results
depend on the shapes at rounding level (the order of the partial sums), so a
tuned run is reproducible given its shapes: ``tune`` returns them and
``save`` / ``load`` keep them in a JSON file keyed by grid, precision and
device, so later runs take the recorded shapes without timing anything.

    from frehg2_amd import autotune
    shapes = autotune.tune(solver)                  # or: autotune.apply(solver, path)
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

from . import abi

CANDIDATES = {
    "march64": (4, 8, 16),
    "march32": (4, 8, 32),
    "column": (8, 16, 32),
}
COUPLING_NAMES = ("eta", "dept", "seep_accum", "avail", "gain")


def _snapshot(g):
    snap = {"h": g.get(abi.F_HEAD), "wc": g.get(abi.F_WC), "vloss": g.get(abi.F_VLOSS)}
    if getattr(g, "_cpl", None) is not None:
        snap["cpl"] = {n: g.coupling_state(n) for n in COUPLING_NAMES}
    return snap


def _restore(g, snap):
    g.set(abi.F_HEAD, snap["h"])
    g.set(abi.F_WC, snap["wc"])
    g.set(abi.F_VLOSS, snap["vloss"])
    for n, v in snap.get("cpl", {}).items():
        g.set_coupling_state(n, v)


def _time_step(g, snap, dt, repeats):
    """Median solve time (HIP events) of one step of dt from snap."""
    times = []
    for _ in range(repeats):
        _restore(g, snap)
        g.synchronize()
        g.set_timing(1)
        g.step(dt, t=g.t)
        g.synchronize()
        times.append(g.timing()["solve_ms"])
    g.set_timing(0)
    return float(np.median(times))


def tune(g, dt=None, repeats=2, candidates=None):
    """Pick the PCG launch shapes of solver ``g`` (frehg2_amd.gw.RichardsSolver)
    that solve one step of ``dt`` (default: its next dt) fastest.  The
    solver's state is restored afterwards; returns the chosen shapes and the
    trial times (ms)."""
    cands = candidates or CANDIDATES
    dt = g.next_dt() if dt is None else dt
    t0 = g.t
    snap = _snapshot(g)
    best = g.pcg_shapes()
    trials = {}

    def key(sh):
        return f"{sh['march64']}/{sh['march32']}/{sh['column']}"

    def measure(sh):
        k = key(sh)
        if k not in trials:
            g.set_pcg_shapes(sh["march64"], sh["march32"], sh["column"])
            trials[k] = _time_step(g, snap, dt, repeats)
            g.t = t0
        return trials[k]

    best_ms = measure(best)
    for knob in ("march64", "march32", "column"):
        for v in cands[knob]:
            sh = dict(best, **{knob: v})
            ms = measure(sh)
            if ms < best_ms:
                best, best_ms = sh, ms
    g.set_pcg_shapes(best["march64"], best["march32"], best["column"])
    _restore(g, snap)
    g.t = t0
    g.synchronize()
    return {"shapes": best, "solve_ms": best_ms, "trials_ms": trials}


COLUMN_TILES = (1, 2, 4, 8)


def tune_columns(g, dt=None, repeats=3, candidates=COLUMN_TILES):
    """Pick the tile (columns per wave) of each depth bucket of the batched
    column solve (include/frehg_gw.h frehg_gw_set_column_tiles) of a
    laterally decoupled solver ``g``: coordinate descent over the buckets,
    deepest first, each trial one step of ``dt`` from the same saved state,
    timed by the step's solve events (median of ``repeats``).  The tile
    changes which wave solves which column, not the arithmetic: the
    solution is bitwise the same for every tile.  Leaves the fastest tiles
    set and the state restored (the dt controller's next value is the trial
    step's); returns them with the trial times (ms)."""
    buckets = g.column_tiles()
    if not buckets:
        raise ValueError("the solver has no column solve (lateral conductivity is nonzero)")
    dt = g.next_dt() if dt is None else dt
    t0 = g.t
    snap = _snapshot(g)
    best = [b["cpw"] for b in buckets]
    trials = {}

    def measure(tiles):
        k = ",".join(map(str, tiles))
        if k not in trials:
            g.set_column_tiles(tiles)
            trials[k] = _time_step(g, snap, dt, repeats)
            g.t = t0
        return trials[k]

    best_ms = measure(best)
    for b in range(len(best)):
        for v in candidates:
            if v > buckets[b]["columns"]:
                continue
            tiles = best[:b] + [v] + best[b + 1:]
            ms = measure(tiles)
            if ms < best_ms:
                best, best_ms = tiles, ms
    g.set_column_tiles(best)
    _restore(g, snap)
    g.t = t0
    g.synchronize()
    return {"tiles": [dict(b, cpw=c) for b, c in zip(buckets, best)], "solve_ms": best_ms,
            "trials_ms": trials}


def cache_key(g, device: str = "") -> str:
    c = g.case
    dev = device or os.environ.get("FREHG_DEVICE_NAME", "mi355x")
    return f"{dev}:{c.nx}x{c.ny}x{c.nz}:{c.precision}:{c.preconditioner}"


def save(path, g, shapes, device: str = "") -> None:
    p = Path(path)
    data = json.loads(p.read_text()) if p.exists() else {}
    data[cache_key(g, device)] = shapes
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


def apply(g, path, device: str = "") -> dict | None:
    """Set the shapes recorded for g's case in ``path``; None if none are."""
    p = Path(path)
    if not p.exists():
        return None
    sh = json.loads(p.read_text()).get(cache_key(g, device))
    if sh:
        g.set_pcg_shapes(sh["march64"], sh["march32"], sh["column"])
    return sh
