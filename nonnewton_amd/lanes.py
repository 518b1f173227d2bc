"""Lanes: the batched form as K pools stepped concurrently.

A pool (nn_pool) steps all of its scenes with one launch sequence per tick on
its own HIP stream.  Inside that sequence the GPU waits where the host reads a
counter to decide what runs next (the ends of the divergence, CG and pressure
loops, the tick boundary), and it runs nearly empty launches while the slowest
scene's CG iterates alone.  A LanePool spreads the resident scenes over K pools
("lanes"), each on its own stream and ticked by its own host thread, with no
barrier between lanes: the lanes drift apart in phase, so one lane's waits and
CG tail overlap the other lanes' full-grid sweeps.  Scenes are independent and
every sum is per scene in a fixed order, so a scene gives the same bits in any
lane and beside any other lanes.

Scenes go to the lane with the smallest load (particles resident, or the cost
the caller gives).  With more than one lane each pool's stream has a hardware
queue of its own (NN_POOL_OWN_QUEUE): the device maps plain streams onto a few
shared queues, and two lanes on one queue would run one after the other.
"""
from __future__ import annotations

import threading
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

Key = Tuple[int, int]  # (lane, slot in that lane's pool)


class LanePool:
    """K pools on their own streams, each ticked by its own host thread."""

    def __init__(self, budget_bytes: int, lanes: int = 3, shapes: Optional[Dict[str, Dict[str, int]]] = None,
                 block: int = 0):
        if lanes < 1:
            raise ValueError("lanes must be at least 1")
        self.budget = int(budget_bytes)
        # each lane may use the whole budget; admit() keeps the sum within it.
        # With several lanes every pool's stream gets a hardware queue of its
        # own: lanes that shared one would run one after the other.
        self.pools: List[_lib.Pool] = []
        try:
            for _ in range(lanes):
                self.pools.append(_lib.Pool(self.budget, own_queue=lanes > 1))
            for p in self.pools:
                if block:
                    p.set_block(block)
                if shapes:
                    from . import autotune

                    autotune.apply(p, shapes)
        except BaseException:  # a pool that failed to be created or configured: release the others
            self.close()
            raise
        self.load = [0.0] * lanes
        self.cost: Dict[Key, float] = {}
        self.bytes: Dict[Key, int] = {}
        # admission (one thread) and retirement (the lanes' threads) update the
        # loads and byte counts under this lock; the pools have their own locks
        self._mu = threading.Lock()
        self._joining: Dict[object, int] = {}  # admissions in progress: their bytes

    @property
    def lanes(self) -> int:
        return len(self.pools)

    def close(self):
        for p in self.pools:
            p.close()

    def footprint(self, params, n: int, n_elastic: int = 0) -> int:
        return self.pools[0].footprint(params, n, n_elastic)

    def bytes_in_use(self) -> int:
        with self._mu:
            return sum(self.bytes.values()) + sum(self._joining.values())

    def admit(self, params, x, kind, v=None, cost: Optional[float] = None, n_elastic: int = 0) -> Optional[Key]:
        """(lane, slot), or None when the budget (over all lanes) cannot hold the
        scene now.  The scene goes to the least loaded lane.  NNError for a
        scene the pool refuses (out of grid, overflow, bad kinds)."""
        n = int(np.asarray(x).reshape(-1, 3).shape[0])
        fp = self.footprint(params, n, n_elastic)
        c = float(n if cost is None else cost)
        token = object()
        with self._mu:
            # a scene is held twice while it joins its lane's batch
            if fp < 0 or sum(self.bytes.values()) + sum(self._joining.values()) + 2 * fp > self.budget:
                return None
            lane = int(np.argmin(self.load))
            self.load[lane] += c
            self._joining[token] = fp
        slot = None
        try:
            slot = self.pools[lane].admit(params, x, kind, v)
        finally:
            with self._mu:
                self._joining.pop(token)
                if slot is None:
                    self.load[lane] -= c
                else:
                    self.cost[(lane, slot)] = c
                    self.bytes[(lane, slot)] = fp
        return None if slot is None else (lane, slot)

    def retire(self, key: Key):
        lane, slot = key
        self.pools[lane].retire(slot)
        with self._mu:
            self.load[lane] -= self.cost.pop(key, 0.0)
            self.bytes.pop(key, None)

    def read(self, key: Key, field: int) -> np.ndarray:
        return self.pools[key[0]].read(key[1], field)

    def stats(self, key: Key) -> dict:
        return self.pools[key[0]].stats(key[1])

    def steps(self, key: Key) -> int:
        return self.pools[key[0]].steps(key[1])

    def tick_all(self):
        """One tick of every lane, the lanes concurrently (a barrier at the end)."""
        self.run(1, [[] for _ in self.pools])

    def run(self, ticks: int, slots: Sequence[Sequence[int]]):
        """`ticks` ticks of every lane, each lane on its own thread with no
        barrier between lanes.  slots[l]: the slots of lane l whose statistics
        are recorded after each of its ticks.  Returns (elapsed seconds, per
        lane: (tick durations [s], list of nn_step_stats arrays))."""
        out: List[Optional[tuple]] = [None] * len(self.pools)
        errors: List[BaseException] = []

        def work(l: int):
            try:
                pool = self.pools[l]
                lat, st = [], []
                for _ in range(ticks):
                    t = time.perf_counter()
                    pool.tick()
                    lat.append(time.perf_counter() - t)
                    if slots[l]:
                        st.append(pool.stats_array(slots[l]))
                out[l] = (lat, st)
            except BaseException as e:  # re-raised by the caller's thread
                errors.append(e)

        threads = [threading.Thread(target=work, args=(l,), daemon=True) for l in range(len(self.pools))]
        t0 = time.perf_counter()
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        elapsed = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return elapsed, out

    # kernel timing over every lane (the lanes' counts and times added)
    def profile(self, enable: bool = True, kernels=None, stride: int = 1):
        for p in self.pools:
            p.profile(enable, kernels=kernels, stride=stride)

    def kernel_launches(self) -> Dict[str, int]:
        tot: Dict[str, int] = {}
        for p in self.pools:
            for k, v in p.kernel_launches().items():
                tot[k] = tot.get(k, 0) + v
        return tot


def split_by_cost(costs: Sequence[float], lanes: int) -> List[List[int]]:
    """Longest-processing-time assignment of items (by cost, largest first) to
    the lane with the least total cost: the admission order LanePool.admit
    follows when the items arrive largest first."""
    out: List[List[int]] = [[] for _ in range(lanes)]
    load = [0.0] * lanes
    for k in sorted(range(len(costs)), key=lambda k: (-costs[k], k)):
        l = int(np.argmin(load))
        out[l].append(k)
        load[l] += costs[k]
    return out
