"""Autotuner of the launch shapes of the step's kernels.

Two shapes per kernel: the workgroup size, and the gather depth (how many
neighbor-list entries a lane loads before using them).  Neither changes a
result: the per-particle arithmetic, the order of every sum and the grouping of
the FP64 partial sums are fixed per wavefront.  So shapes are tuned on the live
workload:
  1. workgroup sizes: one tick per candidate size (all kernels at that size),
     rounds in alternating order, every kernel timed with HIP events; the time
     of a solver kernel is divided by its work in that tick (particles x
     launches, from the scenes' iteration counts), since iteration counts
     drift from tick to tick;
  2. gather depths the same way, with the chosen sizes;
  3. verification: ticks with the chosen shapes against ticks with the
     defaults, interleaved and normalized by the CG work; the defaults stay
     unless the choice is faster.

The choice is cached in a JSON file keyed by device and workload signature
(scene count and fluid particle count to 2 significant digits):
$NONNEWTON_AMD_TUNE_CACHE, or ~/.cache/nonnewton_amd/launch_shapes.json.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Iterable, Optional

from . import _lib

CANDIDATES = (64, 128, 256, 512, 1024)
# kernels with a fixed shape: the scan, and the CG update (workgroups of
# whole waves of one scene)
FIXED = {"scan", "cg_update"}


def _signature(device: str, scenes: int, particles: int) -> str:
    """device | scene count | fluid particle count to 2 significant digits"""
    mag = 10 ** max(0, int(math.log10(max(particles, 1))) - 1)
    return f"{device}|scenes={scenes}|particles~{round(particles / mag) * mag}"


def cache_path() -> str:
    p = os.environ.get("NONNEWTON_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "nonnewton_amd", "launch_shapes.json")


def _load_cache() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(c: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


# gather depths each kernel implements (see nn_pool_set_kernel_depth)
# The CG matvec is not tuned: it runs depths 1, 2, 4 and 8, and with its list
# words requested a batch ahead its default of 2 was the fastest in one pool
# and in lanes, where a depth picked from one pool's per-launch times (1) made
# whole ticks slower.
DEPTHS = {k: (4, 8, 16) for k in ("density_alpha", "div_update", "div_source", "press_source",
                                  "press_update", "nonpressure", "strain")}


def apply(pool, shapes: Dict[str, Dict[str, int]]) -> None:
    """shapes: {kernel: {"block": b, "depth": d}} (0 = default)."""
    for k, sh in shapes.items():
        pool.set_kernel_block(k, int(sh.get("block", 0)))
        if k in DEPTHS:
            pool.set_kernel_depth(k, int(sh.get("depth", 0)))


# solver kernels: launches per tick scale with the iteration counts, so their
# time is compared per unit of work (particles x launches) instead of per tick
_WORK = {
    "cg_matvec": ("iterations_cg", 1), "cg_update": ("iterations_cg", 0),
    "div_update": ("iterations_v", 0), "div_source": ("iterations_v", 0),
    "press_update": ("iterations_p", 0), "press_source": ("iterations_p", 1),
}


def _work(kernel: str, stats, n_fluid) -> float:
    if kernel not in _WORK:
        return 1.0
    key, extra = _WORK[kernel]
    w = sum(n * (st[key] + extra) for st, n in zip(stats, n_fluid))
    return max(float(w), 1.0)


def _tick_cost(pool, slots, n_fluid, kernels_on=True):
    """One tick; returns ({kernel: ms per unit of work}, tick seconds per
    particle-iteration of the CG (the step's dominant, variable work))."""
    import time
    if kernels_on:
        pool.profile(True)
    t = time.perf_counter()
    pool.tick()
    dt = time.perf_counter() - t
    stats = [pool.stats(q) for q in slots]
    out = {}
    if kernels_on:
        for k, (ms, n) in pool.kernel_times().items():
            if n and k not in FIXED:
                out[k] = ms / _work(k, stats, n_fluid)
        pool.profile(False)
    cg = sum(n * (1 + st["iterations_cg"]) for st, n in zip(stats, n_fluid))
    return out, dt / max(cg, 1)


def _measure(pool, slots, n_fluid, setter, cands, rounds, valid=None):
    """Per kernel and candidate: summed work-normalized time over `rounds`
    rounds of one tick per candidate (order alternating per round)."""
    per: Dict[str, Dict[int, float]] = {}
    for rnd in range(rounds):
        for c in (cands if rnd % 2 == 0 else cands[::-1]):
            setter(c)
            cost, _ = _tick_cost(pool, slots, n_fluid)
            for k, v in cost.items():
                if valid is not None and (k not in valid or c not in valid[k]):
                    continue
                per.setdefault(k, {}).setdefault(c, 0.0)
                per[k][c] += v
    return {k: min(t.items(), key=lambda kv: kv[1])[0] for k, t in per.items()}


def tune_pool(pool: "_lib.Pool", slots, n_fluid, rounds: int = 3, candidates: Iterable[int] = CANDIDATES,
              depth_levels: Iterable[int] = (1, 2, 4, 8, 16), use_cache: bool = True,
              verify_rounds: int = 4) -> Dict[str, Dict[str, int]]:
    """Pick a workgroup size and a gather depth per kernel for the pool's
    residents (`slots`, with their fluid particle counts `n_fluid`) and apply
    them.  Advances the pool by rounds * (len(candidates) + len(depth_levels))
    + 2 * verify_rounds steps (ordinary steps of the simulation).  The choice
    is kept only if ticks with it beat ticks with the default shapes
    (interleaved, normalized by CG work); otherwise the defaults stay and {} is
    returned.  Result: {kernel: {"block": b, "depth": d}}, 0 = default."""
    device = _lib.device_name()
    sig = _signature(device, len(slots), int(sum(n_fluid)))
    cache = _load_cache() if use_cache else {}
    if sig in cache:
        shapes = cache[sig]
        apply(pool, shapes)
        return shapes
    names = [k for k in _lib.kernel_names() if k not in FIXED]

    def set_blocks(b):
        for k in names:
            pool.set_kernel_block(k, b)

    def set_depths(level):
        for k, ok in DEPTHS.items():
            pool.set_kernel_depth(k, level if level in ok else 0)

    def set_all(shapes):
        for k in names:
            sh = shapes.get(k, {})
            pool.set_kernel_block(k, sh.get("block", 0))
            if k in DEPTHS:
                pool.set_kernel_depth(k, sh.get("depth", 0))

    best_b = _measure(pool, slots, n_fluid, set_blocks, list(candidates), rounds)
    for k in names:
        pool.set_kernel_block(k, best_b.get(k, 0))
    best_d = _measure(pool, slots, n_fluid, set_depths, list(depth_levels), rounds, valid=DEPTHS)
    shapes = {k: {"block": best_b.get(k, 0), "depth": best_d.get(k, 0)} for k in set(best_b) | set(best_d)}
    # verification: tuned shapes against the defaults, interleaved
    tuned = default = 0.0
    for rnd in range(verify_rounds):
        for use in ((True, False) if rnd % 2 == 0 else (False, True)):
            set_all(shapes if use else {})
            _, c = _tick_cost(pool, slots, n_fluid, kernels_on=False)
            if use:
                tuned += c
            else:
                default += c
    if not tuned < default:
        shapes = {}
    set_all(shapes)
    if use_cache:
        cache[sig] = shapes
        _store_cache(cache)
    return shapes
