"""Autotuner of the tile schedule: picks the trailing-update depth gk
(effective panel width 64*gk) and the grid sizes (persistent, or one
workgroup per output block) for a system
structure by timing candidate schedules on one group of energy points, and
caches the choice in a JSON file keyed by structure, precision and group
size."""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

CACHE = os.environ.get("FDMGPU_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                         "tune_cache.json"))
# trail_wg at least the number of output blocks of a launch gives one
# workgroup per block (the hardware schedules them as slots free up) instead
# of a persistent grid striding over the blocks
ONE_PER_BLOCK = 1 << 24
CANDIDATES = [
    # shallow groups for small systems, whose time is mostly the per-step chain
    {"gk": 4, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 6, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 8, "trail_wg": 1024, "solve_wg": 512, "trail_tile": 64},
    {"gk": 12, "trail_wg": 768, "solve_wg": 512, "trail_tile": 64},
    {"gk": 8, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 12, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 16, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 20, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 24, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64},
    {"gk": 20, "trail_wg": ONE_PER_BLOCK, "solve_wg": 512, "trail_tile": 64, "lookahead": 1},
    {"gk": 12, "trail_wg": 512, "solve_wg": 512, "trail_tile": 128},
]
KNOBS = ("gk", "trail_wg", "solve_wg", "trail_tile", "lookahead")
DEFAULTS = {"trail_tile": 64, "lookahead": 0}


def apply(solver, c: dict) -> None:
    solver.tune(c["gk"], c["trail_wg"], 1, c["solve_wg"], c.get("trail_tile", DEFAULTS["trail_tile"]),
                c.get("lookahead", DEFAULTS["lookahead"]))


def key(lb1, lb2, precision: int, group: int) -> str:
    h = hashlib.sha1(np.ascontiguousarray(lb1, np.int32).tobytes() + np.ascontiguousarray(lb2, np.int32).tobytes())
    return f"{h.hexdigest()[:16]}:p{precision}:g{group}"


def load() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(solver, slots, vals, b, x, lb1, lb2, precision: int, candidates=CANDIDATES, repeats: int = 1,
         log=None) -> dict:
    """Times every candidate schedule on this group (after one warm-up solve
    each, which builds its graphs) and applies and caches the fastest."""
    import torch
    k = key(lb1, lb2, precision, len(slots))
    cache = load()
    if k in cache:
        apply(solver, cache[k]["best"])
        return cache[k]
    results = []
    # the FP64 two-tile update spills (its accumulators do not fit the register file)
    candidates = [c for c in candidates if precision != 0 or c.get("trail_tile", 64) == 64]
    for c in candidates:
        apply(solver, c)
        solver.solve(slots, vals, b, x)  # warm-up: builds this schedule's graphs
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(repeats):
            solver.solve(slots, vals, b, x)
        torch.cuda.synchronize()
        dt = (time.time() - t0) / repeats
        results.append(dict(c, seconds=dt))
        if log:
            log(f"autotune {c}: {dt:.3f}s")
    best = min(results, key=lambda r: r["seconds"])
    apply(solver, best)
    entry = {"best": {k2: best.get(k2, DEFAULTS.get(k2)) for k2 in KNOBS}, "trials": results}
    cache[k] = entry
    try:
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass
    return entry
