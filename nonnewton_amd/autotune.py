"""Autotuner of the launch shapes of the step's kernels.

Every sweep's results are independent of its workgroup size (the per-particle
arithmetic and the grouping of the FP64 partial sums are fixed per wavefront),
so shapes can be tuned on the live workload: the tuner runs a few real ticks of
the pool at each candidate size, times every kernel with HIP events, and keeps
for each kernel the size with the smallest time per tick.  Candidates are tried
in two rounds in opposite orders, so the slow drift of the iteration counts as
the scenes evolve does not favour one size.

The choice is cached in a JSON file keyed by device and workload signature
(scene count and particle count rounded to 2 significant digits):
$NONNEWTON_AMD_TUNE_CACHE, or ~/.cache/nonnewton_amd/launch_shapes.json.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Iterable, Optional

from . import _lib

CANDIDATES = (64, 128, 256, 512, 1024)
# kernels with a fixed shape (scan, the scene finish) or launched once at init
FIXED = {"scan"}


def _signature(device: str, scenes: int, particles: int) -> str:
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


def apply(pool: "_lib.Pool", shapes: Dict[str, int]) -> None:
    for k, b in shapes.items():
        pool.set_kernel_block(k, b)


def tune_pool(pool: "_lib.Pool", scenes: int, particles: int, ticks: int = 2,
              candidates: Iterable[int] = CANDIDATES, use_cache: bool = True) -> Dict[str, int]:
    """Pick a workgroup size per kernel for the pool's current residents and
    apply it.  Advances the pool by 2 * len(candidates) * ticks steps (these are
    ordinary steps of the simulation)."""
    device = _lib.device_name()
    sig = _signature(device, scenes, particles)
    cache = _load_cache() if use_cache else {}
    if sig in cache:
        shapes = {k: int(v) for k, v in cache[sig].items()}
        apply(pool, shapes)
        return shapes
    cands = list(candidates)
    per: Dict[str, Dict[int, float]] = {}
    for order in (cands, cands[::-1]):
        for b in order:
            for k in _lib.kernel_names():
                if k not in FIXED:
                    pool.set_kernel_block(k, b)
            pool.profile(True)
            for _ in range(ticks):
                pool.tick()
            for k, (ms, n) in pool.kernel_times().items():
                if n and k not in FIXED:
                    per.setdefault(k, {}).setdefault(b, 0.0)
                    per[k][b] += ms
            pool.profile(False)
    shapes = {k: min(t.items(), key=lambda kv: kv[1])[0] for k, t in per.items()}
    for k in _lib.kernel_names():  # kernels that did not run keep the default
        if k not in FIXED:
            pool.set_kernel_block(k, 0)
    apply(pool, shapes)
    if use_cache:
        cache[sig] = shapes
        _store_cache(cache)
    return shapes
