"""Serving loop: continuous batching of independent SPH scenes on one GPU.

Requests (a scene and a number of steps) arrive over time.  An admission thread
uploads and initializes each arrived request into the pool as soon as the memory
budget allows (nn_pool_admit runs on its own HIP stream, so it overlaps the tick
running on the pool's stream); the main thread ticks the pool -- one step of
every resident scene per tick -- and retires a scene when its steps are done,
which frees budget for waiting requests.  A scene's results do not depend on
which other scenes share its ticks (every sum is per scene and in a fixed
order), so a request gives the same bits alone and in the mix.

The reference has no serving form; this is the multi-scene driver over the same
step (SPHBase.step, SPH/main.py:915-940).
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _lib


@dataclass
class RequestRecord:
    rid: int
    arrival: float
    steps: int
    n_dynamic: int
    admitted: Optional[float] = None
    finished: Optional[float] = None
    slot: Optional[int] = None
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
    """Continuous batching over an nn_pool with admission by memory budget."""

    def __init__(self, budget_bytes: int = 200 << 30, block: int = 0, keep_results: bool = False):
        self.pool = _lib.Pool(budget_bytes)
        self.budget = int(budget_bytes)
        if block:
            self.pool.set_block(block)
        self.keep_results = keep_results
        self._lock = threading.Lock()
        self._cond = threading.Condition(self._lock)

    def close(self):
        self.pool.close()

    def run(self, requests, time_scale: float = 1.0) -> ServingReport:
        """Serve `requests` (objects with rid, arrival [s], scene, steps).  Arrival
        times are honoured on the wall clock (scaled by time_scale)."""
        recs = [RequestRecord(q.rid, q.arrival * time_scale, q.steps, q.scene.n_dynamic) for q in requests]
        scenes = {q.rid: q.scene for q in requests}
        order = sorted(recs, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        resident: Dict[int, RequestRecord] = {}
        pending: List[RequestRecord] = list(order)
        state = {"stop": False, "freed": 0}

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
                fp = self.pool.footprint(sc.params, sc.n, sc.n_elastic)
                if fp < 0 or 2 * fp > self.budget:
                    reject(nxt, f"needs {2 * fp} bytes to join, over the budget of {self.budget}")
                    continue
                try:
                    slot = self.pool.admit(sc.params, sc.x, sc.kind, sc.v)
                except _lib.NNError as e:  # out of grid, cell/neighbor overflow, bad kinds ...
                    reject(nxt, str(e))
                    continue
                with self._cond:
                    if slot is None:  # over budget: wait for a retirement
                        while state["freed"] == freed_seen and not state["stop"]:
                            if not resident:  # nothing will retire: the request cannot fit now
                                break
                            self._cond.wait(timeout=0.05)
                        if not resident and state["freed"] == freed_seen:
                            reject_locked(nxt, "over the budget with nothing resident")
                        continue
                    pending.pop(0)
                    nxt.slot = slot
                    nxt.admitted = time.perf_counter() - t0
                    resident[slot] = nxt
                    self._cond.notify_all()

        def admitter_guarded():
            try:
                admitter()
            except BaseException as e:  # keep the main loop from waiting on a dead thread
                with self._cond:
                    state["admit_error"] = e
                    self._cond.notify_all()

        th = threading.Thread(target=admitter_guarded, daemon=True)
        th.start()
        tick_ms: List[float] = []
        ticks = 0
        psteps = 0
        peak_res = 0
        peak_bytes = 0
        try:
            while True:
                with self._cond:
                    if "admit_error" in state:
                        raise RuntimeError("admission thread failed") from state["admit_error"]
                    active = {s: r for s, r in resident.items() if r.finished is None}
                    if not active and not pending:
                        break
                    if not active:
                        if not th.is_alive() and pending:
                            raise RuntimeError("admission thread ended with requests pending")
                        self._cond.wait(timeout=0.01)
                        continue
                t = time.perf_counter()
                rc, stepped = self.pool.tick()
                dt = time.perf_counter() - t
                r, q, b = self.pool.info()
                peak_res = max(peak_res, r)
                peak_bytes = max(peak_bytes, b)
                # a scene admitted during the tick joins at the next one: count
                # what the pool actually stepped
                with self._cond:
                    live = [s for s, rec in resident.items() if rec.finished is None]
                done = []
                for s in live:
                    rec = resident[s]
                    n = self.pool.steps(s)
                    if n > rec.done_steps:
                        psteps += (n - rec.done_steps) * rec.n_dynamic
                        rec.done_steps = n
                        if self.pool.stats(s)["status"] & _lib.Pool.STATUS_ERROR_BITS:
                            rec.failed = True  # evicted; the other scenes continue unperturbed
                            done.append(s)
                            continue
                    if rec.done_steps >= rec.steps:
                        done.append(s)
                tick_ms.append(1e3 * dt)
                ticks += 1
                for s in done:
                    rec = resident[s]
                    if self.keep_results:
                        rec.result = {"x": self.pool.read(s, _lib.FIELD_X), "v": self.pool.read(s, _lib.FIELD_V)}
                    rec.finished = time.perf_counter() - t0
                    self.pool.retire(s)
                    with self._cond:
                        del resident[s]
                        state["freed"] += 1
                        self._cond.notify_all()
        finally:
            with self._cond:
                state["stop"] = True
                self._cond.notify_all()
            th.join(timeout=10)
        wall = time.perf_counter() - t0
        lat = [r.finished - r.arrival for r in recs if r.finished is not None]
        wait = [r.admitted - r.arrival for r in recs if r.admitted is not None]
        return ServingReport(wall, ticks, psteps, tick_ms, lat, wait, peak_res, peak_bytes, recs)
