"""Serving loop: continuous batching of independent SPH scenes on one GPU.

Requests (a scene and a number of steps) arrive over time.  An admission thread
uploads and initializes each arrived request into a pool as soon as the memory
budget allows (nn_pool_admit runs on its own HIP stream, so it overlaps the tick
running on the pool's stream); each lane's thread ticks its pool -- one step of
every resident scene of that lane per tick -- and retires a scene when its steps
are done, which frees budget for waiting requests.  A scene's results do not depend on
which other scenes share its ticks (every sum is per scene and in a fixed
order), so a request gives the same bits alone and in the mix.

The reference has no serving form; this is the multi-scene driver over the same
step (SPHBase.step, SPH/main.py:915-940).
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib
from .lanes import Key, LanePool


@dataclass
class RequestRecord:
    rid: int
    arrival: float
    steps: int
    n_dynamic: int
    admitted: Optional[float] = None
    finished: Optional[float] = None
    slot: Optional[Tuple[int, int]] = None  # (lane, slot of that lane's pool)
    done_steps: int = 0
    failed: bool = False
    error: str = ""  # why it failed: rejected at admission or evicted by a step's status
    result: Optional[Dict[str, np.ndarray]] = None


@dataclass
class ServingReport:
    wall_s: float
    ticks: int
    particle_steps: int
    tick_ms: List[float]
    request_latency_s: List[float]
    queue_wait_s: List[float]
    peak_resident: int
    peak_bytes: int
    records: List[RequestRecord] = field(default_factory=list)

    def summary(self) -> dict:
        t = np.asarray(self.tick_ms) if self.tick_ms else np.zeros(1)
        r = np.asarray(self.request_latency_s) if self.request_latency_s else np.zeros(1)
        w = np.asarray(self.queue_wait_s) if self.queue_wait_s else np.zeros(1)
        return {
            "requests": len(self.records),
            "wall_s": round(self.wall_s, 3),
            "ticks": self.ticks,
            "particle_steps_per_s": round(self.particle_steps / self.wall_s, 1) if self.wall_s > 0 else 0.0,
            "p50_step_ms": round(float(np.percentile(t, 50)), 3),
            "p99_step_ms": round(float(np.percentile(t, 99)), 3),
            "p50_request_s": round(float(np.percentile(r, 50)), 3),
            "p99_request_s": round(float(np.percentile(r, 99)), 3),
            "p50_queue_wait_s": round(float(np.percentile(w, 50)), 3),
            "p99_queue_wait_s": round(float(np.percentile(w, 99)), 3),
            "peak_resident_scenes": self.peak_resident,
            "peak_bytes": self.peak_bytes,
        }


class ServingLoop:
    """Continuous batching over lanes of nn_pools with admission by memory budget.

    lanes = 1: one pool, ticked by one thread.  lanes = K: the scenes spread over
    K pools (nonnewton_amd.lanes.LanePool), each ticked by its own thread on its
    own stream with no barrier between them; an arriving request goes to the
    lane with the least work assigned (particles x steps of its resident
    requests).  The budget holds over all lanes together."""

    def __init__(self, budget_bytes: int = 200 << 30, block: int = 0, keep_results: bool = False, lanes: int = 1,
                 shapes=None):
        self.lanes = LanePool(budget_bytes, lanes=lanes, shapes=shapes, block=block)
        self.budget = int(budget_bytes)
        self.keep_results = keep_results
        self._lock = threading.Lock()
        self._cond = threading.Condition(self._lock)

    @property
    def pool(self) -> _lib.Pool:
        """The first lane's pool (the only one with lanes = 1)."""
        return self.lanes.pools[0]

    def close(self):
        self.lanes.close()

    def run(self, requests, time_scale: float = 1.0) -> ServingReport:
        """Serve `requests` (objects with rid, arrival [s], scene, steps).  Arrival
        times are honoured on the wall clock (scaled by time_scale)."""
        recs = [RequestRecord(q.rid, q.arrival * time_scale, q.steps, q.scene.n_dynamic) for q in requests]
        scenes = {q.rid: q.scene for q in requests}
        order = sorted(recs, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        lp = self.lanes
        K = lp.lanes
        resident: Dict[Key, RequestRecord] = {}
        pending: List[RequestRecord] = list(order)
        state = {"stop": False, "freed": 0}
        # peak_res: requests admitted and not yet finished (in HBM, joined or
        # staged); peak_bytes: their footprints (the budget's accounting)
        acc = {"ticks": 0, "psteps": 0, "peak_res": 0, "peak_bytes": 0}
        tick_ms: List[float] = []

        def reject_locked(rec, why):  # with self._cond held
            pending.remove(rec)
            rec.failed = True
            rec.error = why
            rec.finished = time.perf_counter() - t0
            self._cond.notify_all()

        def reject(rec, why):
            with self._cond:
                reject_locked(rec, why)

        def admitter():
            while True:
                with self._cond:
                    if state["stop"]:
                        return
                    if not pending:
                        return
                    nxt = pending[0]
                    now = time.perf_counter() - t0
                    if nxt.arrival > now:
                        self._cond.wait(timeout=min(0.05, nxt.arrival - now))
                        continue
                    freed_seen = state["freed"]
                sc = scenes[nxt.rid]
                # a scene is held twice while it joins the batch: one that
                # needs more than the whole budget that way can never be admitted
                fp = lp.footprint(sc.params, sc.n, sc.n_elastic)
                if fp < 0 or 2 * fp > self.budget:
                    reject(nxt, f"needs {2 * fp} bytes to join, over the budget of {self.budget}")
                    continue
                try:
                    # to the lane with the least work assigned: particles x steps
                    key = lp.admit(sc.params, sc.x, sc.kind, sc.v, cost=float(sc.n) * nxt.steps,
                                   n_elastic=sc.n_elastic)
                except _lib.NNError as e:  # out of grid, cell/neighbor overflow, bad kinds ...
                    reject(nxt, str(e))
                    continue
                with self._cond:
                    if key is None:  # over budget: wait for a retirement
                        while state["freed"] == freed_seen and not state["stop"]:
                            if not resident:  # nothing will retire: the request cannot fit now
                                break
                            self._cond.wait(timeout=0.05)
                        if not resident and state["freed"] == freed_seen:
                            reject_locked(nxt, "over the budget with nothing resident")
                        continue
                    pending.pop(0)
                    nxt.slot = key
                    nxt.admitted = time.perf_counter() - t0
                    resident[key] = nxt
                    acc["peak_res"] = max(acc["peak_res"], len(resident))
                    acc["peak_bytes"] = max(acc["peak_bytes"], lp.bytes_in_use())
                    self._cond.notify_all()

        def guarded(fn, tag):
            def body(*a):
                try:
                    fn(*a)
                except BaseException as e:  # keep the other threads from waiting on a dead one
                    with self._cond:
                        state.setdefault("error", (tag, e))
                        state["stop"] = True
                        self._cond.notify_all()
            return body

        adm = threading.Thread(target=guarded(admitter, "admission"), daemon=True)

        def lane_loop(l: int):
            pool = lp.pools[l]
            while True:
                with self._cond:
                    if state["stop"]:
                        return
                    active = [k for k, r in resident.items() if k[0] == l and r.finished is None]
                    if not active:
                        if not pending and not resident:
                            return  # every request is done
                        if pending and not adm.is_alive():
                            raise RuntimeError("admission thread ended with requests pending")
                        self._cond.wait(timeout=0.01)
                        continue
                t = time.perf_counter()
                pool.tick()
                dt = time.perf_counter() - t
                # a scene admitted during the tick joins at the next one: count
                # what the pool actually stepped
                done = []
                psteps = 0
                for key in active:
                    rec = resident[key]
                    n = pool.steps(key[1])
                    if n > rec.done_steps:
                        psteps += (n - rec.done_steps) * rec.n_dynamic
                        rec.done_steps = n
                        if pool.stats(key[1])["status"] & _lib.Pool.STATUS_ERROR_BITS:
                            rec.failed = True  # evicted; the other scenes continue unperturbed
                            done.append(key)
                            continue
                    if rec.done_steps >= rec.steps:
                        done.append(key)
                with self._cond:
                    tick_ms.append(1e3 * dt)
                    acc["ticks"] += 1
                    acc["psteps"] += psteps
                for key in done:
                    rec = resident[key]
                    if self.keep_results:
                        rec.result = {"x": pool.read(key[1], _lib.FIELD_X), "v": pool.read(key[1], _lib.FIELD_V)}
                    rec.finished = time.perf_counter() - t0
                    lp.retire(key)
                    with self._cond:
                        del resident[key]
                        state["freed"] += 1
                        self._cond.notify_all()

        workers = [threading.Thread(target=guarded(lane_loop, f"lane {l}"), args=(l,), daemon=True) for l in range(K)]
        adm.start()
        for w in workers:
            w.start()
        try:
            for w in workers:
                w.join()
        finally:
            with self._cond:
                state["stop"] = True
                self._cond.notify_all()
            adm.join(timeout=10)
        if "error" in state:
            tag, e = state["error"]
            raise RuntimeError(f"serving {tag} thread failed") from e
        wall = time.perf_counter() - t0
        lat = [r.finished - r.arrival for r in recs if r.finished is not None]
        wait = [r.admitted - r.arrival for r in recs if r.admitted is not None]
        return ServingReport(wall, acc["ticks"], acc["psteps"], tick_ms, lat, wait, acc["peak_res"],
                             acc["peak_bytes"], recs)
