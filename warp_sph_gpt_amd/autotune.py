"""Launch-shape autotuner for the per-subject passes.

The tunable shapes (workgroup size, XCD-aware tile order, porosity
sub-group width; see sphdem_tuning_t) decide only which lanes evaluate which
subjects, so every setting gives the same bits and the tuner may time them
on the live simulation.  For each candidate it enqueues a few real steps
with per-pass HIP-event timing and keeps, pass by pass, the fastest shape.
The porosity's waves per work item (fast FP32 path) regroup its fixed-point
partial sums, within that path's tolerance; they are timed the same way,
together with the order in which the waves take the work items.  Where the
fused step runs its DEM interpolation, wall pass and force pass (beside
other passes on library streams, or in sequence) changes only the overlap;
each is chosen on the median of single-step times of the two settings.

    shapes = autotune(engine.step, steps=3)
"""
from __future__ import annotations

import json
import os
import statistics

import torch

from . import _lib

BLOCKS = (64, 128, 256)
REMAPS = (0, 1)
X_TILES = (3, 4, 5)
POR_PARTS = (2, 4, 8, 16)
POR_QUEUES = (0, 1)


def _median_step_ms(step_fn, steps: int, knob: str) -> dict:
    """Median milliseconds of one whole step (every pass, all streams: the
    side streams join back before a step ends) with `knob` at 1 and at 0.
    Each setting runs four times, alternating (1 0 0 1 1 0 0 1), for one
    settling step and steps + 2 timed ones; the median over a setting's
    timed steps leaves out the odd costlier step (a Shepard refresh)."""
    m = steps + 2
    times = {1: [], 0: []}
    for b in (1, 0, 0, 1, 1, 0, 0, 1):
        _lib.set_tuning({knob: b})
        step_fn()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(m + 1)]
        ev[0].record()
        for i in range(m):
            step_fn()
            ev[i + 1].record()
        torch.cuda.synchronize()
        times[b] += [ev[i].elapsed_time(ev[i + 1]) for i in range(m)]
    return {b: sorted(t)[len(t) // 2] for b, t in times.items()}


def _time_steps(step_fn, steps: int) -> dict:
    """Per-pass milliseconds of `steps` steps: `steps` times the median over
    the steps of each pass's time, so that the odd costlier step (a Shepard
    refresh) does not decide a comparison (with 3 or more steps)."""
    step_fn()                      # settle the new shapes (first-launch costs)
    torch.cuda.synchronize()
    per = []
    for _ in range(steps):
        with _lib.PassTimer() as t:
            step_fn()
            torch.cuda.synchronize()
        per.append(t.ms)
    return {k: steps * statistics.median(p.get(k, 0.0) for p in per)
            for k in set().union(*per)}


def autotune(step_fn, steps: int = 3, passes=_lib.TUNABLE_PASSES, cache: str | None = None,
             key: str | None = None, log=None) -> dict:
    """Choose block size and tile order per pass, then the porosity's x tile,
    its waves per work item and how the waves take the items, then where the
    DEM interpolation, the wall pass and the force pass run (each on whole
    steps).  Leaves the library set to the result and returns it (with the
    timings).  With `cache` and `key`, a stored result for `key` is applied
    instead."""
    if cache and key and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            _lib.set_tuning(stored[key]["tuning"])
            return stored[key]
    # the per-pass timings run the force pass alone (beside the porosity the
    # two share the chip); where it runs is chosen on whole steps at the end
    force_beside = _lib.get_tuning()["force_beside"]
    _lib.set_tuning({"force_beside": 0})
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
    # where the DEM interpolation, the wall pass and the force pass run
    # changes only the overlap: whole steps, each switch on its own
    whole = {}
    for knob, needs, what in (
            ("dem_beside", ("porosity",), "the DEM interpolation beside / after the porosity"),
            ("wall_beside", ("wall_density",), "the wall pass beside / after the fluid density"),
            ("force_beside", ("porosity", "fluid_force"),
             "the force pass beside / before the porosity")):
        if not all(p in passes for p in needs):
            continue
        whole[knob] = _median_step_ms(step_fn, steps, knob)
        cfg[knob] = min(whole[knob], key=whole[knob].get)
        _lib.set_tuning({knob: cfg[knob]})
        if log:
            log(f"[autotune] median step ms with {what}: "
                f"{whole[knob][1]:.3f} / {whole[knob][0]:.3f}")
    if "force_beside" not in whole:
        _lib.set_tuning({"force_beside": force_beside})
    out = {"tuning": cfg, "ms": table, "por_x_tile_ms": {str(k): v for k, v in xt_ms.items()},
           "por_queue_remap_parts_ms": {f"{q}/{r}/{pp}": v
                                        for (q, r, pp), v in parts_ms.items()},
           **{f"{k}_median_step_ms": {str(b): v for b, v in ms.items()}
              for k, ms in whole.items()},
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
    taken from one counter, the wall pass beside the fluid density, the
    force pass beside the porosity)."""
    _lib.set_tuning({"block": {p: 256 for p in _lib.PASS_NAMES},
                     "xcd_remap": {p: 0 for p in _lib.PASS_NAMES}, "por_x_tile": 4,
                     "dem_beside": 1, "por_parts": 4, "por_queue": 1, "wall_beside": 1,
                     "force_beside": 1})
