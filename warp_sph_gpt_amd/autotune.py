"""Launch-shape autotuner for the per-subject passes.

The tunable shapes (workgroup size, XCD-aware tile order, porosity
sub-group width; see sphdem_tuning_t) decide only which lanes evaluate which
subjects, so every setting gives the same bits and the tuner may time them
on the live simulation.  For each candidate it enqueues a few real steps
with per-pass HIP-event timing and keeps, pass by pass, the fastest shape.
The porosity's waves per work item (fast FP32 path) regroup its fixed-point
partial sums, within that path's tolerance; they are timed the same way,
together with the order in which the waves take the work items.

    shapes = autotune(engine.step, steps=3)
"""
from __future__ import annotations

import json
import os

import torch

from . import _lib

BLOCKS = (64, 128, 256)
REMAPS = (0, 1)
X_TILES = (3, 4, 5)
POR_PARTS = (2, 4, 8, 16)
POR_QUEUES = (0, 1)


def _time_whole_steps(step_fn, steps: int) -> float:
    """Milliseconds of `steps` steps on the current stream (every pass, both
    streams: the side stream joins back before a step ends)."""
    step_fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        step_fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b)


def _time_steps(step_fn, steps: int) -> dict:
    step_fn()                      # settle the new shapes (first-launch costs)
    torch.cuda.synchronize()
    with _lib.PassTimer() as t:
        for _ in range(steps):
            step_fn()
        torch.cuda.synchronize()
    return dict(t.ms)


def autotune(step_fn, steps: int = 3, passes=_lib.TUNABLE_PASSES, cache: str | None = None,
             key: str | None = None, log=None) -> dict:
    """Choose block size and tile order per pass, then the porosity x tile.
    Leaves the library set to the result and returns it (with the timings).
    With `cache` and `key`, a stored result for `key` is applied instead."""
    if cache and key and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            _lib.set_tuning(stored[key]["tuning"])
            return stored[key]
    table = {p: {} for p in passes}
    for b in BLOCKS:
        for r in REMAPS:
            _lib.set_tuning({"block": {p: b for p in passes}, "xcd_remap": {p: r for p in passes}})
            ms = _time_steps(step_fn, steps)
            for p in passes:
                table[p][f"{b}/{r}"] = ms.get(p, 0.0)
            if log:
                log(f"[autotune] block {b} remap {r}: "
                    + ", ".join(f"{p}={ms.get(p, 0.0) / steps:.3f}" for p in passes))
    best = {p: min(table[p], key=table[p].get) for p in passes}
    cfg = {"block": {p: int(best[p].split("/")[0]) for p in passes},
           "xcd_remap": {p: int(best[p].split("/")[1]) for p in passes}}
    _lib.set_tuning(cfg)
    xt_ms = {}
    if "porosity" in passes:
        for xt in X_TILES:
            _lib.set_tuning({"por_x_tile": xt})
            xt_ms[xt] = _time_steps(step_fn, steps).get("porosity", 0.0)
        cfg["por_x_tile"] = min(xt_ms, key=xt_ms.get)
        _lib.set_tuning({"por_x_tile": cfg["por_x_tile"]})
    parts_ms = {}
    if "porosity" in passes:
        # waves per work item, how the waves take the items and (which only
        # matters when they take them by wave index) the tile order, jointly
        for q in POR_QUEUES:
            for r in (REMAPS if q == 0 else (cfg["xcd_remap"]["porosity"],)):
                for pp in POR_PARTS:
                    _lib.set_tuning({"por_parts": pp, "por_queue": q,
                                     "xcd_remap": {"porosity": r}})
                    parts_ms[(q, r, pp)] = _time_steps(step_fn, steps).get("porosity", 0.0)
        q, r, pp = min(parts_ms, key=parts_ms.get)
        cfg["por_queue"], cfg["por_parts"] = q, pp
        cfg["xcd_remap"]["porosity"] = r
        _lib.set_tuning({"por_parts": pp, "por_queue": q, "xcd_remap": {"porosity": r}})
        if log:
            log("[autotune] porosity ms by (queue, remap, parts): " + ", ".join(
                f"{k}={v / steps:.3f}" for k, v in parts_ms.items()))
    # where the DEM interpolation and the wall pass run changes only the
    # overlap: whole steps
    beside_ms, wall_ms = {}, {}
    if "porosity" in passes:
        for b in (1, 0):
            _lib.set_tuning({"dem_beside": b})
            beside_ms[b] = _time_whole_steps(step_fn, steps)
        cfg["dem_beside"] = min(beside_ms, key=beside_ms.get)
        _lib.set_tuning({"dem_beside": cfg["dem_beside"]})
        if log:
            log(f"[autotune] step ms with the DEM interpolation beside / after the porosity: "
                f"{beside_ms[1] / steps:.3f} / {beside_ms[0] / steps:.3f}")
    if "wall_density" in passes:
        for b in (1, 0):
            _lib.set_tuning({"wall_beside": b})
            wall_ms[b] = _time_whole_steps(step_fn, steps)
        cfg["wall_beside"] = min(wall_ms, key=wall_ms.get)
        _lib.set_tuning({"wall_beside": cfg["wall_beside"]})
        if log:
            log(f"[autotune] step ms with the wall pass beside / after the fluid density: "
                f"{wall_ms[1] / steps:.3f} / {wall_ms[0] / steps:.3f}")
    out = {"tuning": cfg, "ms": table, "por_x_tile_ms": {str(k): v for k, v in xt_ms.items()},
           "por_queue_remap_parts_ms": {f"{q}/{r}/{pp}": v
                                        for (q, r, pp), v in parts_ms.items()},
           "dem_beside_ms": {str(k): v for k, v in beside_ms.items()},
           "wall_beside_ms": {str(k): v for k, v in wall_ms.items()},
           "steps_per_candidate": steps}
    if log:
        log(f"[autotune] chosen {json.dumps(cfg)}")
    if cache and key:
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[key] = out
        with open(cache, "w") as f:
            json.dump(stored, f, indent=1)
    return out


def reset() -> None:
    """Back to the library defaults (256 threads, no remap, x tile 4, the
    DEM interpolation beside the porosity, 4 porosity waves per work item
    taken from one counter, the wall pass beside the fluid density)."""
    _lib.set_tuning({"block": {p: 256 for p in _lib.PASS_NAMES},
                     "xcd_remap": {p: 0 for p in _lib.PASS_NAMES}, "por_x_tile": 4,
                     "dem_beside": 1, "por_parts": 4, "por_queue": 1, "wall_beside": 1})
