"""Serving loop on the GPU: requests admitted while others run give the same bits
as when run alone; a failing scene is evicted without perturbing the others; a
tight memory budget queues requests until retirements free it.  Each case runs
with one pool and with three lanes."""
import numpy as np
import pytest

from nonnewton_amd import _lib as L
from nonnewton_amd import scenes
from nonnewton_amd.scenes import Request
from nonnewton_amd.serving import ServingLoop

pytestmark = pytest.mark.gpu


def alone(sc, steps):
    c = L.Context(sc.params, sc.n)
    c.upload(sc.x, sc.kind, sc.v)
    c.init()
    for _ in range(steps):
        c.step()
    out = {"x": c.download(L.FIELD_X), "v": c.download(L.FIELD_V)}
    c.close()
    return out


def requests(steps=(3, 5, 2, 4), seed=3):
    batch = scenes.mixed_batch(seed=seed, count=len(steps), lo=2000, hi=4000)
    return [Request(k, 0.02 * k, sc, st) for k, (sc, st) in enumerate(zip(batch, steps))]


@pytest.mark.parametrize("lanes", [1, 3])
def test_serving_results_equal_alone(lanes):
    reqs = requests()
    loop = ServingLoop(8 << 30, keep_results=True, lanes=lanes)
    rep = loop.run(reqs)
    loop.close()
    s = rep.summary()
    assert s["requests"] == 4 and all(r.finished is not None for r in rep.records)
    for rec, q in zip(rep.records, reqs):
        ref = alone(q.scene, q.steps)
        assert rec.done_steps == q.steps
        assert np.array_equal(rec.result["x"], ref["x"]) and np.array_equal(rec.result["v"], ref["v"])
    assert s["p99_step_ms"] >= s["p50_step_ms"] > 0


@pytest.mark.parametrize("lanes", [1, 3])
def test_serving_evicts_failing_scene(lanes):
    reqs = requests(steps=(3, 3, 3))
    bad = reqs[1].scene
    v = bad.v.copy()
    v[0] = np.nan  # forced from the host: the scene's state turns non-finite
    bad.v = v
    loop = ServingLoop(8 << 30, keep_results=True, lanes=lanes)
    rep = loop.run(reqs)
    loop.close()
    assert rep.records[1].failed and rep.records[1].done_steps == 1
    for k in (0, 2):
        ref = alone(reqs[k].scene, 3)
        assert not rep.records[k].failed
        assert np.array_equal(rep.records[k].result["x"], ref["x"])


@pytest.mark.parametrize("lanes", [1, 3])
def test_serving_budget_queues_requests(lanes):
    reqs = requests(steps=(2, 2, 2, 2), seed=5)
    pool = L.Pool(1 << 34)
    fp = max(pool.footprint(q.scene.params, q.scene.n, q.scene.n_elastic) for q in reqs)
    pool.close()
    loop = ServingLoop(int(3.5 * fp), keep_results=True, lanes=lanes)
    rep = loop.run(reqs)
    loop.close()
    assert rep.peak_resident <= 2
    assert all(r.finished is not None and not r.failed for r in rep.records)
    for rec, q in zip(rep.records, reqs):
        assert np.array_equal(rec.result["x"], alone(q.scene, q.steps)["x"])


@pytest.mark.parametrize("lanes", [1, 3])
def test_serving_rejects_what_cannot_be_admitted(lanes):
    """A request larger than the budget allows and one with a particle outside
    its grid are marked failed with a reason; the loop returns and the other
    requests are served as when alone."""
    reqs = requests(steps=(2, 2, 2, 2), seed=6)
    pool = L.Pool(1 << 34)
    fp = max(pool.footprint(q.scene.params, q.scene.n, q.scene.n_elastic) for q in reqs)
    pool.close()
    big = scenes.mixed_scene(np.random.Generator(np.random.PCG64(9)), 60000, 0, 1.0, 1.0, False, "big")
    reqs[1] = Request(1, reqs[1].arrival, big, 2)
    bad = reqs[2].scene
    x = bad.x.copy()
    x[0] = (-1.0, 0.1, 0.1)  # below the domain: out of grid at admission
    bad.x = x
    loop = ServingLoop(int(2.5 * fp), keep_results=True, lanes=lanes)
    rep = loop.run(reqs)
    loop.close()
    r = rep.records
    assert r[1].failed and "budget" in r[1].error and r[1].slot is None
    assert r[2].failed and r[2].error and r[2].slot is None
    for k in (0, 3):
        assert not r[k].failed and r[k].done_steps == 2
        assert np.array_equal(r[k].result["x"], alone(reqs[k].scene, 2)["x"])


def test_serving_in_lanes_equals_alone():
    """Three lanes (pools on their own streams, each ticked by its own thread):
    every request gives the bits it gives alone, and the lanes all take part."""
    reqs = requests(steps=(3, 5, 2, 4, 3, 2), seed=7)
    loop = ServingLoop(8 << 30, keep_results=True, lanes=3)
    rep = loop.run(reqs, time_scale=0.0)  # all arrive at once, so they overlap
    loop.close()
    assert all(r.finished is not None and not r.failed for r in rep.records)
    assert {r.slot[0] for r in rep.records} == {0, 1, 2}
    for rec, q in zip(rep.records, reqs):
        ref = alone(q.scene, q.steps)
        assert rec.done_steps == q.steps
        assert np.array_equal(rec.result["x"], ref["x"]) and np.array_equal(rec.result["v"], ref["v"])
