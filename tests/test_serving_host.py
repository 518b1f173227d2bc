"""Host logic of the serving loop and the lanes, on the CPU: the pools are
replaced by a fake with nn_pool's contract (a scene admitted during a tick
joins at the next one; a footprint per scene; a budget per pool; status bits
after a step), so admission by budget, queueing, rejection, eviction, the peak
accounting and the spread over lanes run without a GPU."""
import threading
from types import SimpleNamespace

import numpy as np
import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd import lanes as lanes_mod
from nonnewton_amd.scenes import Request
from nonnewton_amd.serving import ServingLoop

BYTES_PER_PARTICLE = 1000


class FakePool:
    """nn_pool's behaviour as the serving loop sees it."""

    STATUS_ERROR_BITS = L.Pool.STATUS_ERROR_BITS

    def __init__(self, budget_bytes, own_queue=False):
        self.budget = budget_bytes
        self.own_queue = own_queue
        self.mu = threading.Lock()
        self.next_slot = 0
        self.slots = {}  # slot -> dict(n, steps, joined, bad, status, x)
        self.ticks = 0

    def close(self):
        pass

    def set_block(self, block):
        pass

    def footprint(self, params, n, n_elastic=0):
        return BYTES_PER_PARTICLE * int(n)

    def admit(self, params, x, kind, v=None):
        x = np.asarray(x, dtype=np.float32).reshape(-1, 3)
        if not np.isfinite(x).all():
            raise L.NNError(L.NN_E_OUT_OF_GRID, "nn_pool_admit", "particle outside its grid")
        if getattr(params, "explode", False):
            raise ValueError("admission failed for a reason that is not a scene error")
        fp = self.footprint(params, len(x))
        with self.mu:
            used = sum(self.footprint(None, s["n"]) for s in self.slots.values())
            if used + 2 * fp > self.budget:
                return None
            slot = self.next_slot
            self.next_slot += 1
            self.slots[slot] = {"n": len(x), "steps": 0, "joined": False, "bad": getattr(params, "bad", False),
                                "status": 0, "x": x.copy()}
            return slot

    def tick(self):
        with self.mu:
            self.ticks += 1
            for s in self.slots.values():
                if not s["joined"]:
                    s["joined"] = True  # joins at this tick's start: stepped from the next one
                    continue
                s["steps"] += 1
                if s["bad"]:
                    s["status"] |= 8  # non-finite state
        return 0, len(self.slots)

    def steps(self, slot):
        with self.mu:
            return self.slots[slot]["steps"]

    def stats(self, slot):
        with self.mu:
            return {"status": self.slots[slot]["status"]}

    def read(self, slot, field):
        with self.mu:
            return self.slots[slot]["x"] + self.slots[slot]["steps"]  # the state after that many steps

    def retire(self, slot):
        with self.mu:
            del self.slots[slot]


@pytest.fixture
def fake_pools(monkeypatch):
    monkeypatch.setattr(lanes_mod._lib, "Pool", FakePool)


def scene(n, seed=0, bad=False, explode=False, finite=True):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)).astype(np.float32)
    if not finite:
        x[0, 0] = np.nan
    params = SimpleNamespace(bad=bad, explode=explode)
    return SimpleNamespace(params=params, x=x, kind=np.zeros(n, np.int32), v=np.zeros((n, 3), np.float32),
                           n=n, n_dynamic=n, n_elastic=0)


def requests(sizes, steps, spacing=0.0):
    return [Request(k, spacing * k, scene(n, seed=k), st) for k, (n, st) in enumerate(zip(sizes, steps))]


@pytest.mark.parametrize("lanes", [1, 3])
def test_every_request_served_with_its_steps(fake_pools, lanes):
    reqs = requests([100, 200, 300, 400, 500, 600], [3, 5, 2, 4, 6, 1])
    loop = ServingLoop(1 << 30, keep_results=True, lanes=lanes)
    rep = loop.run(reqs, time_scale=0.0)
    loop.close()
    for rec, q in zip(rep.records, reqs):
        assert rec.finished is not None and not rec.failed
        assert rec.done_steps == q.steps
        np.testing.assert_array_equal(rec.result["x"], q.scene.x + q.steps)
    assert rep.particle_steps == sum(n * st for n, st in zip([100, 200, 300, 400, 500, 600], [3, 5, 2, 4, 6, 1]))
    assert rep.peak_resident >= 1 and rep.peak_bytes >= BYTES_PER_PARTICLE * 100
    if lanes == 3:
        assert {r.slot[0] for r in rep.records} == {0, 1, 2}


@pytest.mark.parametrize("lanes", [1, 3])
def test_budget_queues_requests(fake_pools, lanes):
    """Each scene needs twice its footprint while it joins: with 3.5 footprints
    of budget at most two are resident at once, and every request is served."""
    n = 100
    reqs = requests([n] * 6, [2] * 6)
    loop = ServingLoop(int(3.5 * BYTES_PER_PARTICLE * n), keep_results=True, lanes=lanes)
    rep = loop.run(reqs, time_scale=0.0)
    loop.close()
    assert 1 <= rep.peak_resident <= 2
    assert rep.peak_bytes <= 3.5 * BYTES_PER_PARTICLE * n
    assert all(r.finished is not None and not r.failed and r.done_steps == 2 for r in rep.records)
    waits = sorted(r.admitted for r in rep.records)
    assert waits[-1] > waits[0]  # later requests waited for retirements


@pytest.mark.parametrize("lanes", [1, 3])
def test_rejections_and_eviction(fake_pools, lanes):
    """Too large for the budget and out of grid: failed at admission with a
    reason; a scene whose state turns bad: evicted after its first step; the
    others are served."""
    reqs = requests([100, 100, 100, 100], [3, 3, 3, 3])
    reqs[1] = Request(1, 0.0, scene(10000), 3)  # 2 x 10 MB > the 5 MB budget
    reqs[2] = Request(2, 0.0, scene(100, finite=False), 3)
    reqs[3] = Request(3, 0.0, scene(100, bad=True), 3)
    loop = ServingLoop(5 * BYTES_PER_PARTICLE * 1000, keep_results=True, lanes=lanes)
    rep = loop.run(reqs, time_scale=0.0)
    loop.close()
    r = rep.records
    assert not r[0].failed and r[0].done_steps == 3
    assert r[1].failed and "budget" in r[1].error and r[1].slot is None
    assert r[2].failed and "NN_E_OUT_OF_GRID" in r[2].error and r[2].slot is None
    assert r[3].failed and r[3].done_steps == 1 and r[3].slot is not None


def test_admission_thread_failure_is_raised(fake_pools):
    """An unexpected admission error ends the loop with an error instead of a
    wait that never ends."""
    reqs = requests([100, 100], [50, 2])
    reqs[1] = Request(1, 0.0, scene(100, explode=True), 2)
    loop = ServingLoop(1 << 30, lanes=2)
    with pytest.raises(RuntimeError, match="admission"):
        loop.run(reqs, time_scale=0.0)
    loop.close()


def test_lane_pool_budget_over_all_lanes(fake_pools):
    """LanePool keeps the sum of the lanes' footprints (twice a joining
    scene's) within one budget and sends each scene to the least loaded lane."""
    lp = lanes_mod.LanePool(10 * BYTES_PER_PARTICLE * 100, lanes=3)
    keys = [lp.admit(None, scene(100, seed=k).x, None) for k in range(8)]
    assert keys[:8] == [(0, 0), (1, 0), (2, 0), (0, 1), (1, 1), (2, 1), (0, 2), (1, 2)]
    assert lp.bytes_in_use() == 8 * BYTES_PER_PARTICLE * 100
    assert lp.admit(None, scene(100).x, None) == (2, 2)  # 8 + 2 x 1 footprints: exactly the budget
    assert lp.admit(None, scene(100).x, None) is None  # 9 + 2 x 1 > 10
    lp.retire((0, 0))
    assert lp.bytes_in_use() == 8 * BYTES_PER_PARTICLE * 100
    assert lp.admit(None, scene(100).x, None) == (0, 3)  # lane 0 now holds the fewest particles
    lp.close()
